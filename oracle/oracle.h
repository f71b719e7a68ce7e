/*
 * oracle.h — CPU restatement of the falx reference path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or as the timed CPU baseline. The product (libfalx_gpu.so) never
 * links, loads or calls it.
 *
 * `threads <= 1` runs the reference's serial algorithm; larger values run
 * the reference's parallel driver for the dialect (parity prepass with a
 * barrier, speculative index with redo, region transfer functions, or
 * line-aligned chunking), whose results are identical to serial.
 */
#ifndef FALX_ORACLE_H
#define FALX_ORACLE_H

#include <stdint.h>
#include "../include/falx_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* scalar.rs:23-60, byte at a time. Returns the number of offsets. */
uint64_t ora_index_scalar(const falx_spec* spec, const uint8_t* data, uint64_t len,
                          uint32_t* out);
/* Generated-kernel semantics: 64-byte blocks with carries (interp.rs,
 * formats.rs:206-297, csv_hash.rs:2124-2220). `threads` selects the
 * parallel driver (csv_geo.rs:95-149 or csv_hash.rs:345-400). */
uint64_t ora_index_structurals(const falx_spec* spec, const uint8_t* data,
                               uint64_t len, int threads, uint32_t* out);

/* parse / parse_par tape (csv_geo.rs:394-421, 215-334, 1838-1849). */
void ora_parse(const falx_spec* spec, const uint8_t* data, uint64_t len, int threads,
               uint32_t* seps, uint64_t* n_seps, uint64_t* ends, uint64_t* n_ends);

/* parse_columns / parse_columns_par into host buffers sized for
 * rows_cap rows (rows_cap >= len + 1 always suffices) and data_cap >= len. */
int ora_parse_columns(const falx_spec* spec, const uint8_t* data, uint64_t len,
                      int threads, falx_columns* out);

/* Fused sinks. The f64 `sum` fields are sequential left folds. */
void ora_csv_stats(const falx_spec* spec, const uint8_t* data, uint64_t len,
                   int threads, falx_csv_stats* out);
void ora_field_bytes(const falx_spec* spec, const uint8_t* data, uint64_t len,
                     int threads, falx_field_byte_stats* out);
uint64_t ora_ndjson_lines(const uint8_t* data, uint64_t len, int threads);
void ora_ndjson_id_score(const uint8_t* data, uint64_t len, int threads,
                         falx_ndjson_id_score_stats* out);
void ora_logfmt_pairs(const uint8_t* data, uint64_t len, int threads,
                      falx_logfmt_stats* out);

/* Cell parsers (codegen.rs:7500-7744 templates). Return 1 when valid. */
int ora_parse_f64(const uint8_t* s, uint64_t n, double* out, int* fixed6,
                  int64_t* mantissa);
int ora_parse_i64(const uint8_t* s, uint64_t n, int64_t* out);
/* Dialect cleaning (codegen.rs:8008-8185). Returns the cleaned length. */
uint64_t ora_clean(const falx_spec* spec, const uint8_t* s, uint64_t n, uint8_t* out);

/* json::parse_nested / parse_nested_par (kernels/json.rs:636-760,
 * 996-1108): the nested tape of the JSON structural index into tape[0, cap)
 * (cap >= len always suffices); returns its length. error / error_pos as
 * falx_parse_nested. */
uint64_t ora_parse_nested(const uint8_t* data, uint64_t len, int threads, uint64_t* tape,
                          uint64_t cap, int32_t* error, uint64_t* error_pos);

/* Raw DEFLATE (RFC 1951) into out[0, cap): the output length, -1 when
 * malformed (the reference's flate2/libdeflater backends, restated). */
int64_t ora_inflate(const uint8_t* src, uint64_t n, uint8_t* out, uint64_t cap);
/* bgzf::decompress_par over a falx_bgzf_scan block table: the total, or -1. */
int64_t ora_bgzf_inflate(const uint8_t* comp, const falx_bgzf_block* blocks, uint64_t n_blocks,
                         int threads, uint8_t* out);

/* Which block-classifier tier the CPU dispatches ("avx512bw" or "scalar"). */
const char* ora_isa(void);

#ifdef __cplusplus
}
#endif

#endif
