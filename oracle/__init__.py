"""CPU oracle for the falx structural-indexing / column-parsing path.

Test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg). The product package falx_amd never imports it.
"""
