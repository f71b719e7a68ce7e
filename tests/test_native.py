"""CPU sweeps of the device-side code paths that are written once for host
and device (falx_amd/csrc/cells.h, ndjson_walk.h, num_swar.h, regions.h):
their host build is run against the oracle on randomized cells and
records, and the block region resolver against a byte-at-a-time automaton."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")


def _build(name: str, ora) -> str:
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, name)
    lib = os.path.join(ROOT, "oracle", "_build")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "native", name + ".cpp"),
                    os.path.join(lib, "liboracle.so"), f"-Wl,-rpath,{lib}"], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 0xDEADBEEF, 20261203])
def test_cell_parsers_match_oracle(ora, seed):
    exe = _build("cells_check", ora)
    r = subprocess.run([exe, str(seed), "150000", "15000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "agree" in r.stdout


@pytest.mark.parametrize("seed", [1, 77])
def test_ndjson_walker_matches_oracle(ora, seed):
    exe = _build("ndjson_check", ora)
    r = subprocess.run([exe, str(seed), "100000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("seed", [3, 0xC0FFEE])
def test_register_number_parser_matches_oracle(ora, seed):
    exe = _build("swar_check", ora)
    r = subprocess.run([exe, str(seed), "300000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "agree" in r.stdout


def test_fixed6_divide_free_quotient_exhaustive(ora):
    """div_1e6, the fixed-6 cells' m / 10^6 without a divide, equals the IEEE
    quotient for every mantissa below 10^9."""
    exe = _build("fixed6_div_check", ora)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "agree" in r.stdout


@pytest.mark.parametrize("seed", [5, 0xFEED])
def test_block_regions_match_byte_automaton(ora, seed):
    exe = _build("regions_check", ora)
    r = subprocess.run([exe, str(seed), "400000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "agree" in r.stdout


@pytest.mark.parametrize("seed", [5, 0xD0B1E])
def test_doubled_quote_drops_match_stream(ora, seed):
    """dq_dropped (cells.h), the doubled-quote drops the string passes read
    from a tile's quote bitmap, against DoubledStream on random insides."""
    exe = _build("dq_check", ora)
    r = subprocess.run([exe, str(seed), "200000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "agree" in r.stdout
