/*
 * falx_gpu.h — C ABI of the MI355X (gfx950) structural-indexing and
 * column-parsing library.
 *
 * The reference exposes this path as the public functions of a generated
 * Rust module (one module per format spec, src/kernels/<fmt>.rs). This
 * header is the device backend a `falx::gpu` shim binds with `extern "C"`:
 * plain pointers and sizes, status codes instead of panics, no unwinding.
 *
 * Conventions
 *  - Every `d_` pointer is device memory (HBM). Inputs stay read-only.
 *  - Entry points run on the context's stream and return after the result
 *    counts are known (they synchronise that stream once per call).
 *  - Offsets are u32, so inputs must be shorter than 2^32 bytes; longer
 *    inputs return FALX_E_TOO_LARGE (the reference casts `len as u32`
 *    unchecked, src/kernels/csv_geo.rs:1113).
 *  - Output capacities are caller-provided. When a result does not fit,
 *    the call returns FALX_E_CAPACITY and still reports the required size
 *    in the count out-parameters, so the caller can grow and retry.
 *  - A context is bound to one device and one stream; calls on one
 *    context must not overlap (use one context per host thread).
 */
#ifndef FALX_GPU_H
#define FALX_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum falx_status {
    FALX_OK = 0,
    FALX_E_INVALID = 1,    /* bad argument or spec (spec.rs:44-276 rules) */
    FALX_E_TOO_LARGE = 2,  /* input >= 2^32 bytes, or a string column
                              beyond the 2 GiB Arrow i32-offset limit
                              (reference panics, csv_hash.rs:1405-1408) */
    FALX_E_CAPACITY = 3,   /* an output buffer is too small; counts report
                              the required sizes */
    FALX_E_DEVICE = 4,     /* HIP runtime error */
    FALX_E_NO_DEVICE = 5,  /* no usable gfx950 device */
    FALX_E_SCHEMA = 6,     /* an NDJSON record lacks the schema keys
                              (reference panics, ndjson.rs:720-723) */
    FALX_E_BUSY = 7        /* admission refused: HBM budget exhausted */
} falx_status;

#define FALX_MAX_STRUCTURAL 16
#define FALX_MAX_COLUMNS 32

/* Column cell types (codegen.rs:78-94). */
typedef enum falx_column_type {
    FALX_COL_I64 = 0,   /* exactly `str::parse::<i64>` */
    FALX_COL_F64 = 1,   /* exactly `str::parse::<f64>` (correctly rounded) */
    FALX_COL_STR = 2,   /* cleaned bytes, Arrow varbinary (offsets + data) */
    FALX_COL_BYTES = 3  /* raw (start, end) span into the input */
} falx_column_type;

typedef struct falx_column {
    uint32_t index;     /* zero-based field ordinal within a record */
    uint32_t type;      /* falx_column_type */
} falx_column;

/*
 * Runtime dialect + projection descriptor. It carries what the reference
 * bakes into a generated module at build time: formats.rs:30-53 (Dialect)
 * and codegen.rs:99-110 (Column). `structural` lists field separators and,
 * normally, '\n' (the record terminator class, formats.rs:206-224).
 */
typedef struct falx_spec {
    uint8_t structural[FALX_MAX_STRUCTURAL];
    uint32_t n_structural;
    int32_t quote;      /* -1 = no quote convention */
    int32_t escape;     /* -1 = none / doubled quotes; else backslash-style
                           escape byte (formats.rs:15-25) */
    int32_t comment;    /* -1 = none; else line-start comment byte */
    uint32_t n_columns;
    falx_column columns[FALX_MAX_COLUMNS];
} falx_spec;

/* One projected column's device buffers (Arrow layouts, csv_typed.rs:920-977). */
typedef struct falx_column_buf {
    void* values;       /* I64: int64[rows_cap]; F64: double[rows_cap];
                           BYTES: uint32[2*rows_cap] (start, end);
                           STR: int32 offsets[rows_cap + 1] */
    uint64_t* valid;    /* LSB-first validity, ceil(rows_cap / 64) words */
    uint8_t* data;      /* STR only: cleaned bytes arena */
    uint64_t data_cap;  /* STR only: arena capacity (bytes) */
    uint64_t data_len;  /* STR only, out: arena bytes used */
} falx_column_buf;

typedef struct falx_columns {
    uint64_t rows_cap;  /* in: row capacity of every column buffer */
    uint64_t rows;      /* out: rows produced (required rows on CAPACITY) */
    falx_column_buf cols[FALX_MAX_COLUMNS];
} falx_columns;

/*
 * Fused per-column statistics. `values`/`checksum`/`bytes` are the
 * reference's fused counters generalised to any column set:
 *  F64: values = parsed cells, checksum = wrapping sum of f64 bit patterns
 *       (csv_geo.rs:1735-1750);
 *  STR: values = present cells, bytes = cleaned length, checksum = wrapping
 *       sum of cleaned bytes (csv_geo_text.rs:1870-1878);
 *  I64: values = parsed cells, checksum = wrapping sum of the values;
 *  BYTES: values = non-empty cells, bytes = raw length, checksum = byte sum.
 * Added on top: `sum` (F64/I64 values, blocked FP64 tree) and, for F64
 * cells of the fixed `[-]d{1,3}.d{6}` shape, the exact int64 sum of their
 * signed 1e-6 mantissas.
 */
typedef struct falx_column_stats {
    uint64_t values;
    uint64_t checksum;
    uint64_t bytes;
    int64_t fixed6_sum;
    uint64_t fixed6_cells;
    double sum;
} falx_column_stats;

typedef struct falx_csv_stats {
    uint64_t records;   /* every non-comment record, header included */
    falx_column_stats cols[FALX_MAX_COLUMNS];
} falx_csv_stats;

typedef struct falx_field_byte_stats {
    uint64_t fields;
    uint64_t bytes;
} falx_field_byte_stats;

typedef struct falx_logfmt_stats {
    uint64_t pairs;
    uint64_t key_bytes;
    uint64_t value_bytes;
    uint64_t checksum;
} falx_logfmt_stats;

/*
 * NDJSON `sum(id + nested.score)`. `id_sum` is the wrapping i64 sum of the
 * top-level integer `id`; scores are JSON numbers parsed correctly rounded
 * to f64 and summed with a blocked FP64 tree (`score_sum`); scores with at
 * most six decimals and no exponent are also summed exactly in 1e-6 units
 * (`score_fixed6_sum`, over `score_fixed6_cells` records). `invalid` counts
 * non-blank records without both keys (the reference panics on those).
 */
typedef struct falx_ndjson_id_score_stats {
    uint64_t records;
    uint64_t invalid;
    int64_t id_sum;
    int64_t score_fixed6_sum;
    uint64_t score_fixed6_cells;
    double score_sum;
} falx_ndjson_id_score_stats;

typedef struct falx_ctx falx_ctx;

/* ---- context, memory ---------------------------------------------------- */

/* Create a context on `device` (HIP ordinal) with its own stream. */
falx_status falx_ctx_create(int device, falx_ctx** out);
void falx_ctx_destroy(falx_ctx* ctx);
/* The context's hipStream_t, for callers that order their own work. */
void* falx_ctx_stream(falx_ctx* ctx);
/* Device milliseconds of the last entry point's main kernel(s): HIP events
 * on the context stream around the scan launches (fills and final
 * reductions excluded). */
double falx_ctx_last_kernel_ms(falx_ctx* ctx);
/* Detail of the last FALX_E_DEVICE from an entry point (no counterpart in the
 * reference, whose CPU path has no device to fail): the HIP error code of the
 * failing runtime call on the calling thread (0 = none). */
falx_status falx_ctx_last_error(const falx_ctx* ctx, int32_t* hip_error);

falx_status falx_device_alloc(falx_ctx* ctx, uint64_t bytes, void** d_out);
falx_status falx_device_free(falx_ctx* ctx, void* d_ptr);
falx_status falx_host_alloc_pinned(falx_ctx* ctx, uint64_t bytes, void** h_out);
falx_status falx_host_free_pinned(falx_ctx* ctx, void* h_ptr);
falx_status falx_upload(falx_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes);
falx_status falx_download(falx_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes);

/* spec.rs:44-276 validation rules applied to a runtime spec. */
falx_status falx_spec_validate(const falx_spec* spec);

/* ---- structural index ----------------------------------------------------- */

/*
 * Byte offsets of every structural byte (src/kernels/csv_geo.rs:24-52
 * `index_structurals`, and `index_structurals_par` 95-149 / csv_hash.rs:345-400,
 * which are byte-identical). Writes n_out offsets to d_out.
 */
falx_status falx_index_structurals(falx_ctx* ctx, const falx_spec* spec,
                                   const uint8_t* d_data, uint64_t len,
                                   uint32_t* d_out, uint64_t out_cap,
                                   uint64_t* n_out);

/*
 * falx_index_structurals without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, outputs and statuses as
 * falx_index_structurals, but the call's scratch (tile summaries, prefixes,
 * result slots) lives in a file-scope device array of the library, about
 * 2 MiB per device, held until the process ends; it makes no hipMalloc,
 * hipFree or other allocating call (the context is created beforehand).
 * Inputs longer than FALX_STATIC_MAX_LEN return FALX_E_TOO_LARGE. All calls
 * on one device share that array: calls that overlap on one GPU (from
 * several contexts or threads) race and must be serialised by the caller.
 */
#define FALX_STATIC_MAX_LEN (64ull << 20)
falx_status falx_index_structurals_static(falx_ctx* ctx, const falx_spec* spec,
                                          const uint8_t* d_data, uint64_t len,
                                          uint32_t* d_out, uint64_t out_cap,
                                          uint64_t* n_out);

/* ---- tape ----------------------------------------------------------------- */

/*
 * The record tape of `parse` / `parse_par` (csv_geo.rs:394-421, 215-334,
 * push_tape 1838-1849): separator offsets, and per terminated record an end
 * entry `(separators before it << 32) | terminator offset`.
 */
falx_status falx_parse(falx_ctx* ctx, const falx_spec* spec,
                       const uint8_t* d_data, uint64_t len,
                       uint32_t* d_seps, uint64_t seps_cap,
                       uint64_t* d_ends, uint64_t ends_cap,
                       uint64_t* n_seps, uint64_t* n_ends);

/* ---- typed columns -------------------------------------------------------- */

/*
 * Arrow-layout typed columns of the spec's projection: `parse_columns` /
 * `parse_columns_par` (csv_geo.rs:1112-1116, 1216-1284; csv_hash.rs:1570-1663;
 * csv_typed.rs:979-1152). Column buffers must be sized for `rows_cap` rows;
 * validity words are fully written (bits past `rows` are zero).
 */
falx_status falx_parse_columns(falx_ctx* ctx, const falx_spec* spec,
                               const uint8_t* d_data, uint64_t len,
                               falx_columns* out);

/*
 * Typed columns plus the fused statistics of the same columns from one pass
 * over the input: the reference's column lane followed by the sums over the
 * materialised columns (examples/bench_columns.rs:151-179,
 * examples/csv_geo_aggregate.rs:42-63), with `stats` as falx_parse_csv_stats
 * would return it (counts, checksums and fixed-6 sums equal; FP64 sums are
 * fixed-order trees that may differ from it in the last bits). On
 * FALX_E_CAPACITY the statistics of the numeric and bytes columns cover the
 * rows that fit in `rows_cap` only.
 */
falx_status falx_parse_columns_with_stats(falx_ctx* ctx, const falx_spec* spec,
                                          const uint8_t* d_data, uint64_t len,
                                          falx_columns* out, falx_csv_stats* stats);

/* ---- fused sinks ---------------------------------------------------------- */

/*
 * Fused count/checksum/sum over the spec's columns without materialising
 * them: `parse_csv_geo_stats[_par]` (csv_geo.rs:1505-1770) and
 * `parse_csv_geo_text_stats[_par]` (csv_geo_text.rs:1609-1913), with the
 * column ordinals taken from the spec instead of hardcoded.
 */
falx_status falx_parse_csv_stats(falx_ctx* ctx, const falx_spec* spec,
                                 const uint8_t* d_data, uint64_t len,
                                 falx_csv_stats* out);

/* `parse_field_bytes[_par]` (csv_hash.rs:965-1227; tsv.rs:847-866). */
falx_status falx_parse_field_bytes(falx_ctx* ctx, const falx_spec* spec,
                                   const uint8_t* d_data, uint64_t len,
                                   falx_field_byte_stats* out);

/* `parse_ndjson_lines[_par]` (ndjson.rs:618-657): escape-aware line count. */
falx_status falx_parse_ndjson_lines(falx_ctx* ctx, const uint8_t* d_data,
                                    uint64_t len, uint64_t* records);

/* `parse_ndjson_id_score[_par]` (ndjson.rs:679-749), schema-aware form. */
falx_status falx_parse_ndjson_id_score(falx_ctx* ctx, const uint8_t* d_data,
                                       uint64_t len,
                                       falx_ndjson_id_score_stats* out);

/* `parse_logfmt_pairs[_par]` (logfmt.rs:612-857). */
falx_status falx_parse_logfmt_pairs(falx_ctx* ctx, const uint8_t* d_data,
                                    uint64_t len, falx_logfmt_stats* out);

/* ---- JSON nested tape ------------------------------------------------------
 * `json::parse_nested[_par]` (kernels/json.rs:636-656, 996-1108) and the
 * `Nested` tape it returns (json.rs:617-630): one entry per
 * structural byte outside strings (`{ } [ ] , :`, formats.rs:164-173), in
 * input order. A bracket's entry is (partner's tape index << 32) | byte
 * position, the partner being 0xFFFFFFFF for an open never closed; a `,` or
 * `:` entry is its position. `error` is NestError (json.rs:608-615): on a
 * close with no open or of the wrong kind the tape stops before it
 * (*n_tape = its tape index) and error = FALX_NEST_UNMATCHED_CLOSE at its
 * position; input ending with brackets open gives FALX_NEST_UNCLOSED_OPEN at
 * the innermost open's position. FALX_E_CAPACITY when the tape needs more
 * than tape_cap entries (*n_tape = the count needed). */
typedef enum falx_nest_error {
    FALX_NEST_OK = 0,
    FALX_NEST_UNMATCHED_CLOSE = 1,
    FALX_NEST_UNCLOSED_OPEN = 2,
} falx_nest_error;

falx_status falx_parse_nested(falx_ctx* ctx, const uint8_t* d_data, uint64_t len,
                              uint64_t* d_tape, uint64_t tape_cap, uint64_t* n_tape,
                              int32_t* error, uint64_t* error_pos);

/* ---- streaming ingestion ---------------------------------------------------
 * `StreamParser` (csv_geo.rs:688-807 and every generated module's
 * `stream()`): feed chunks of unbounded input, receive the records each feed
 * completes. The unfinished last record stays in a device buffer of
 * `buf_cap` bytes (a record plus a chunk must fit) and is indexed again with
 * the next chunk behind it, so quotes, escapes and comments across feed
 * boundaries parse exactly as in one document.
 *
 * Each call returns the tape of the records it completes in the caller's
 * arrays, over `out->d_buf` (device; valid until the next call on the
 * stream): ends[i] = (separators before it << 32) | terminator offset, the
 * first record starting at offset 0 and record i + 1 after terminator i,
 * seps the separators of those records. `data_end` is where the emitted
 * records stop: the byte after the last terminator (feed), or the end of
 * the buffer (finish, whose last record is unterminated when data_end is
 * past the last terminator). Records are skipped as the tape iterators skip
 * them (comment records). FALX_E_CAPACITY leaves the stream unchanged:
 * `need_bytes` is the buffer size the feed needs when the kept record and
 * the chunk do not fit (0 otherwise), else the counts are the tape sizes
 * needed. */
typedef struct falx_stream falx_stream;
typedef struct falx_stream_out {
    const uint8_t* d_buf;
    uint64_t n_seps;
    uint64_t n_ends;
    uint64_t data_end;
    uint64_t need_bytes;
} falx_stream_out;

falx_status falx_stream_create(falx_ctx* ctx, const falx_spec* spec, uint64_t buf_cap,
                               falx_stream** out);
void falx_stream_destroy(falx_stream* stream);
/* Start a new stream on the same buffers (no allocation). */
falx_status falx_stream_reset(falx_stream* stream);
falx_status falx_stream_feed(falx_stream* stream, const uint8_t* d_chunk, uint64_t len,
                             uint32_t* d_seps, uint64_t seps_cap, uint64_t* d_ends,
                             uint64_t ends_cap, falx_stream_out* out);
falx_status falx_stream_finish(falx_stream* stream, uint32_t* d_seps, uint64_t seps_cap,
                               uint64_t* d_ends, uint64_t ends_cap, falx_stream_out* out);

/* ---- BGZF ------------------------------------------------------------------
 * `bgzf::scan` / `bgzf::decompress_par` (bgzf.rs:61-290): a BGZF stream is a
 * sequence of gzip members of at most 64 KiB output each, independently
 * inflatable, with the payload length in the `BC` extra subfield and the
 * inflated size (ISIZE) in the trailer. falx_bgzf_scan walks the headers of
 * host bytes into a block table (empty EOF blocks dropped; `out` is each
 * block's offset in the inflated stream, *total_out their sum);
 * FALX_E_INVALID for a malformed header (the reference's InvalidData),
 * FALX_E_CAPACITY when more than `cap` blocks (*n_blocks = the count).
 * falx_bgzf_inflate inflates those blocks (raw DEFLATE, RFC 1951) from the
 * compressed bytes in HBM into d_out, one wave per block; a malformed block
 * or one whose inflated size differs from ISIZE gives FALX_E_INVALID and
 * *first_bad its index (the smallest). The parsers then read d_out. */
typedef struct falx_bgzf_block {
    uint64_t payload;      /* raw-DEFLATE payload offset in the compressed bytes */
    uint32_t payload_len;
    uint32_t isize;        /* inflated size */
    uint64_t out;          /* offset of its output */
} falx_bgzf_block;

falx_status falx_bgzf_scan(const uint8_t* h_comp, uint64_t len, falx_bgzf_block* blocks,
                           uint64_t cap, uint64_t* n_blocks, uint64_t* total_out);
falx_status falx_bgzf_inflate(falx_ctx* ctx, const uint8_t* d_comp, uint64_t comp_len,
                              const falx_bgzf_block* blocks, uint64_t n_blocks,
                              uint8_t* d_out, uint64_t out_cap, uint64_t* first_bad);

/* ---------------------------------------------------------------------
 * Arrow export. The reference hands its columns to Arrow by wrapping the
 * buffers, since they are Arrow's own layouts (examples/arrow_interop.rs:
 * 1-115). Here the columns live in HBM, so a column is exported through the
 * Arrow C Device Data Interface (the public ABI below, as the Arrow project
 * specifies it) with device_type ARROW_DEVICE_ROCM: the array's buffers are
 * the column's device buffers themselves, nothing is copied.
 * ------------------------------------------------------------------- */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4
struct ArrowSchema {
    const char* format;
    const char* name;
    const char* metadata;
    int64_t flags;
    int64_t n_children;
    struct ArrowSchema** children;
    struct ArrowSchema* dictionary;
    void (*release)(struct ArrowSchema*);
    void* private_data;
};
struct ArrowArray {
    int64_t length;
    int64_t null_count;
    int64_t offset;
    int64_t n_buffers;
    int64_t n_children;
    const void** buffers;
    struct ArrowArray** children;
    struct ArrowArray* dictionary;
    void (*release)(struct ArrowArray*);
    void* private_data;
};
#endif
#ifndef ARROW_C_DEVICE_DATA_INTERFACE
#define ARROW_C_DEVICE_DATA_INTERFACE
typedef int32_t ArrowDeviceType;
#define ARROW_DEVICE_CPU 1
#define ARROW_DEVICE_ROCM 10
struct ArrowDeviceArray {
    struct ArrowArray array;
    int64_t device_id;
    ArrowDeviceType device_type;
    void* sync_event;
    int64_t reserved[3];
};
#endif

/*
 * Column k of a columns result (falx_parse_columns[_with_stats] or a batch)
 * as an Arrow array of out->rows rows, zero-copy over its buffers:
 *   I64 -> int64 ("l"), F64 -> float64 ("g"): validity, values;
 *   STR -> binary ("z"): validity, int32 offsets, data (the reference's
 *          BinaryArray, arrow_interop.rs:67-74);
 *   BYTES -> fixed-size list of 2 uint32 ("+w:2" of "I"): the (start, end)
 *          span of each cell into the input.
 * null_count is -1 (not counted: the bitmap stays on the device). The
 * arrays do not own the buffers: their release callbacks free only what
 * the export allocated on the host, so the column buffers must outlive
 * them. device_type is `device_type` (ARROW_DEVICE_ROCM for HBM buffers,
 * ARROW_DEVICE_CPU for columns in host memory), sync_event is NULL (the
 * entry points return after their stream has finished).
 */
falx_status falx_export_arrow(const falx_spec* spec, const falx_columns* cols, uint32_t k,
                              int32_t device_type, int64_t device_id,
                              struct ArrowDeviceArray* out, struct ArrowSchema* schema);

/* ---------------------------------------------------------------------
 * Batches. No counterpart in the reference, which parses one document per
 * call (the generated module functions above). A batch runs many documents
 * of uneven sizes and mixed specs through each pipeline stage in one
 * persistent launch over a device table of (document, chunk of tiles)
 * units, which workgroups take by stride; the launch sequence is captured
 * in a hipGraph per set of stages
 * and replayed. Every document's outputs and result are bit-identical to
 * the plain entry point's (falx_index_structurals, falx_parse_columns[_with_stats],
 * falx_parse_csv_stats, falx_parse_field_bytes, falx_parse_logfmt_pairs,
 * falx_parse_ndjson_id_score). One batch in flight per falx_batch; use one
 * falx_batch (and falx_ctx) per concurrent slot.
 * ------------------------------------------------------------------- */
typedef enum falx_op {
    FALX_OP_INDEX = 1,            /* falx_index_structurals */
    FALX_OP_COLUMNS = 2,          /* falx_parse_columns */
    FALX_OP_COLUMNS_STATS = 3,    /* falx_parse_columns_with_stats */
    FALX_OP_CSV_STATS = 4,        /* falx_parse_csv_stats */
    FALX_OP_FIELD_BYTES = 5,      /* falx_parse_field_bytes */
    FALX_OP_LOGFMT_PAIRS = 6,     /* falx_parse_logfmt_pairs (no spec) */
    FALX_OP_NDJSON_ID_SCORE = 7   /* falx_parse_ndjson_id_score (no spec) */
} falx_op;

typedef struct falx_request {
    uint32_t op;                  /* falx_op */
    uint32_t chunk;               /* tiles per unit; 0 = the tuning's choice */
    const falx_spec* spec;        /* ops 1-5; read at submit */
    const uint8_t* d_data;        /* device input, 16-byte aligned */
    uint64_t len;
    falx_columns* columns;        /* COLUMNS*: device buffers; rows and data_len
                                     are written by falx_batch_wait */
    uint32_t* d_index;            /* INDEX: output offsets */
    uint64_t index_cap;
} falx_request;

typedef struct falx_result {
    int32_t status;               /* the plain entry point's status; FALX_E_BUSY
                                     when the batch's scratch could not hold it */
    uint32_t reserved;
    uint64_t n_index;             /* INDEX */
    falx_csv_stats stats;         /* COLUMNS_STATS, CSV_STATS */
    falx_field_byte_stats field_bytes;
    falx_logfmt_stats logfmt;
    falx_ndjson_id_score_stats ndjson;
} falx_result;

/* Launch shape. chunk[k]: tiles per unit for documents of [4^k,
 * 4^(k+1)) tiles (8 KiB each); workgroups: persistent grid of the tile
 * stages and of the other stages. Chosen per spec and size class by
 * falx_amd/tune.py. */
typedef struct falx_batch_tuning {
    uint32_t tile_workgroups;
    uint32_t unit_workgroups;
    uint32_t chunk[8];
    uint32_t use_graph;           /* 0: launch the stages directly; k >= 1: batches of at least k
                                     documents capture their launch sequence in a hipGraph (once
                                     per set of stages) and replay it */
} falx_batch_tuning;

typedef struct falx_batch falx_batch;

/* scratch_bytes: device memory for the per-document scratch of one batch
 * (see falx_request_scratch_bytes); admission control sizes batches by it. */
falx_status falx_batch_create(falx_ctx* ctx, uint64_t scratch_bytes, falx_batch** out);
void falx_batch_destroy(falx_batch* batch);
uint64_t falx_request_scratch_bytes(const falx_request* req);
falx_status falx_batch_get_tuning(const falx_batch* batch, falx_batch_tuning* out);
falx_status falx_batch_set_tuning(falx_batch* batch, const falx_batch_tuning* tuning);
/* Enqueues the batch on the context's stream and returns. */
falx_status falx_batch_submit(falx_batch* batch, const falx_request* reqs, uint32_t n);
/* FALX_OK when the batch in flight (if any) has finished, FALX_E_BUSY if not. */
falx_status falx_batch_query(falx_batch* batch);
/* Waits and fills results[0..n) of the last submit. */
falx_status falx_batch_wait(falx_batch* batch, falx_result* results);
/* Device time of the last batch: descriptor upload to results download. */
double falx_batch_last_ms(const falx_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* FALX_GPU_H */
