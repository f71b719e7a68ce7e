// falx_gpu.hip — C ABI of include/falx_gpu.h over the gfx950 kernels in
// scan.hip. One context = one device + one HIP stream + reusable scratch
// (tile summaries and prefixes, per-tile partials, result slots).
#include "scan.hip"
#include "nested.h"
#include "inflate.h"

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <new>

using namespace fx;

struct falx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_count = nullptr;   // a count copied home mid-call
    float last_ms = 0.f;
    unsigned long long* d_result = nullptr;
    unsigned long long* h_result = nullptr;
    void* d_lb = nullptr;
    size_t lb_cap = 0;
    void* d_part = nullptr;
    size_t part_cap = 0;
    void* d_red = nullptr;       // reduction outputs (StatPart[32] + extras)
    unsigned char* h_red = nullptr;   // pinned home of the reduction outputs a call reads
    bool timed = false;
};

namespace {

constexpr size_t RED_BYTES = 4 << 20;   // reduction scratch: level-1 partials, then results
constexpr size_t RED_OUT = 2 << 20;     // offset of the level-2 results inside it
constexpr size_t REC_OFF = 64 * 1024;   // offset of the record count behind the StatParts (from RED_OUT)
constexpr size_t H_RED_BYTES = 4096;    // the StatParts and the record count, copied home together

inline falx_status hip_status(hipError_t e) { return e == hipSuccess ? FALX_OK : FALX_E_DEVICE; }

// HIP error of the last failing runtime call on this thread.
thread_local int32_t tl_hip_error = 0;

#define HIP_TRY(x)                                     \
    do {                                               \
        hipError_t _e = (x);                           \
        if (_e != hipSuccess) {                        \
            tl_hip_error = (int32_t)_e;                \
            return FALX_E_DEVICE;                      \
        }                                              \
    } while (0)

constexpr size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Entry states of each dialect kind (see dialect_kind): quote x2, backslash
// escape x2, comment regions x3.
constexpr uint32_t kKindStates[6] = {1, 3, 2, 4, 3, 6};

constexpr uint32_t summary_stride(uint32_t ns) { return (uint32_t)TS_COUNTS + (uint32_t)C7 * ns; }

// Per-context scratch for a launch over n tiles: tile summaries, group
// summaries, group and tile prefixes, and per tile and string column the
// cleaned byte count and arena offset.
struct LbLayout {
    size_t tsum, gsum, gpre, tpre, str_len, str_base, fix, total;
};

constexpr LbLayout lb_layout(size_t n, uint32_t n_str, uint32_t ns) {
    const size_t stride = 4 * (size_t)summary_stride(ns);
    const size_t groups = (n + SG - 1) / SG;
    LbLayout L{};
    size_t o = 0;
    L.tsum = o; o = align_up(o + stride * n, 256);
    L.gsum = o; o = align_up(o + stride * groups, 256);
    L.gpre = o; o = align_up(o + 32 * groups, 256);
    L.tpre = o; o = align_up(o + 32 * n, 256);
    L.str_len = o; o = align_up(o + 4 * (size_t)n_str * n, 256);
    L.str_base = o; o = align_up(o + 4 * (size_t)n_str * n, 256);
    L.fix = o; o = align_up(o + 4 * n, 256);
    L.total = o;
    return L;
}

falx_status ensure_lb(falx_ctx* c, uint32_t n_tiles, uint32_t n_str, uint32_t ns) {
    LbLayout L = lb_layout(n_tiles, n_str, ns);
    if (L.total <= c->lb_cap) return FALX_OK;
    if (c->d_lb) HIP_TRY(hipFree(c->d_lb));
    c->d_lb = nullptr;
    c->lb_cap = 0;
    size_t cap = align_up(L.total + L.total / 4, 1 << 20);
    HIP_TRY(hipMalloc(&c->d_lb, cap));
    HIP_TRY(hipMemsetAsync(c->d_lb, 0, cap, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->lb_cap = cap;
    return FALX_OK;
}

falx_status ensure_part(falx_ctx* c, size_t bytes) {
    if (bytes <= c->part_cap) return FALX_OK;
    if (c->d_part) HIP_TRY(hipFree(c->d_part));
    c->d_part = nullptr;
    c->part_cap = 0;
    size_t cap = align_up(bytes + bytes / 4, 1 << 20);
    HIP_TRY(hipMalloc(&c->d_part, cap));
    c->part_cap = cap;
    return FALX_OK;
}

int dialect_kind(const falx_spec* s) {
    bool hq = s->quote >= 0, esc = hq && s->escape >= 0, cmt = s->comment >= 0;
    if (!hq) return cmt ? 1 : 0;
    if (!cmt) return esc ? 3 : 2;
    return esc ? 5 : 4;
}

// Launches one k_scan pass over all tiles, a workgroup per tile.
template <int MODE>
hipError_t launch_scan(falx_ctx* c, Params& P, int kind) {
    return launch_tile_kernel(kind, F_PLAIN, MODE, dim3(P.n_tiles), c->stream, &P, nullptr, 0);
}

// The full form of a pass over the tiles its fast form flagged.
template <int MODE>
hipError_t launch_redo(falx_ctx* c, Params& P, int kind) {
    return launch_tile_kernel(kind, F_REDO, MODE, dim3((P.n_tiles + 63) / 64), c->stream, &P, nullptr, 0);
}

// A COLUMNS or STATS pass as its fast form over all tiles, then the full
// form over the tiles the fast one flagged (cells outside the register
// parsers' reach: quoted or long numbers, inf/nan, undecided roundings).
template <int FAST, int FULL>
hipError_t launch_fast(falx_ctx* c, Params& P, int kind) {
    const LbLayout L = lb_layout(P.n_tiles, P.n_str, kKindStates[kind]);
    P.fixup = reinterpret_cast<uint32_t*>(static_cast<char*>(c->d_lb) + L.fix);
    hipError_t e = launch_scan<FAST>(c, P, kind);
    if (e == hipSuccess) e = launch_redo<FULL>(c, P, kind);
    P.fixup = nullptr;
    return e;
}

// Dialects with neither escape nor comment (kinds 0 and 2) have numeric fast
// forms, for passes whose columns [first, first + n) are all i64/f64.
bool numeric_fast(const falx_spec* spec, int kind, uint32_t first, uint32_t n) {
    if (kind != 0 && kind != 2) return false;
    for (uint32_t k = first; k < first + n && k < spec->n_columns; ++k) {
        const uint32_t t = spec->columns[k].type;
        if (t != FALX_COL_I64 && t != FALX_COL_F64) return false;
    }
    return true;
}

// Whether the statistics of a columns call need the pass over columns
// [g, g + STAT_GROUP): the columns pass fuses those of the first
// COL_STAT_GROUP columns that are not strings, the STRINGS pass those of
// every string column; a group holding any other column needs its pass.
bool stats_group_needed(const falx_spec* spec, uint32_t g) {
    uint32_t rank = 0;   // columns before k that are not strings
    for (uint32_t k = 0; k < g + STAT_GROUP && k < spec->n_columns; ++k) {
        if (spec->columns[k].type == FALX_COL_STR) continue;
        if (k >= g && rank >= (uint32_t)COL_STAT_GROUP) return true;
        ++rank;
    }
    return false;
}

// The statistics pass over columns [P.col_group, + STAT_GROUP).
hipError_t launch_stats_group(falx_ctx* c, const falx_spec* spec, Params& P, int kind) {
    if (numeric_fast(spec, kind, P.col_group, STAT_GROUP)) return launch_fast<M_NUM_STATS_FAST, M_STATS>(c, P, kind);
    return launch_fast<M_STATS_FAST, M_STATS>(c, P, kind);
}

// Levels 2 and 3 of the tile-prefix scan for NS entry states.
template <int NS>
void launch_scan_par(falx_ctx* c, const uint32_t* tsum, const uint32_t* gsum, uint32_t stride, uint32_t n_tiles,
                     uint32_t groups, uint32_t* gpre, uint32_t* tpre, unsigned long long* result) {
    hipLaunchKernelGGL(k_scan_groups_par<NS>, dim3(1), dim3(GB_THREADS), 0, c->stream, gsum, stride, groups, gpre,
                       result);
    hipLaunchKernelGGL(k_scan_tiles_par<NS>, dim3(groups), dim3(SG), 0, c->stream, tsum, stride, n_tiles, gpre, tpre);
}

// The tile-prefix pipeline: tile summaries (every entry state), then the
// three-level scan to each tile's entry state and exclusive counters, and
// the grand totals in the result slots. Emit passes read P.tpre.
// `base` is the scratch laid out by lb_layout (the context's, by default).
hipError_t run_prefix(falx_ctx* c, Params& P, int kind, char* base = nullptr) {
    const uint32_t ns = kKindStates[kind];
    const uint32_t stride = summary_stride(ns);
    const uint32_t groups = (P.n_tiles + SG - 1) / SG;
    LbLayout L = lb_layout(P.n_tiles, P.n_str, ns);
    if (!base) base = static_cast<char*>(c->d_lb);
    uint32_t* tsum = reinterpret_cast<uint32_t*>(base + L.tsum);
    uint32_t* gsum = reinterpret_cast<uint32_t*>(base + L.gsum);
    uint32_t* gpre = reinterpret_cast<uint32_t*>(base + L.gpre);
    uint32_t* tpre = reinterpret_cast<uint32_t*>(base + L.tpre);
    P.tsum = tsum;
    P.ts_stride = stride;
    P.tpre = nullptr;
    hipError_t e = launch_scan<M_SUMMARY>(c, P, kind);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_reduce, dim3(groups), dim3(SG), 0, c->stream, tsum, stride, P.n_tiles, ns, gsum);
    switch (ns) {
    case 1: launch_scan_par<1>(c, tsum, gsum, stride, P.n_tiles, groups, gpre, tpre, P.result); break;
    case 2: launch_scan_par<2>(c, tsum, gsum, stride, P.n_tiles, groups, gpre, tpre, P.result); break;
    case 3: launch_scan_par<3>(c, tsum, gsum, stride, P.n_tiles, groups, gpre, tpre, P.result); break;
    case 4: launch_scan_par<4>(c, tsum, gsum, stride, P.n_tiles, groups, gpre, tpre, P.result); break;
    default: launch_scan_par<6>(c, tsum, gsum, stride, P.n_tiles, groups, gpre, tpre, P.result); break;
    }
    P.tpre = tpre;
    return hipGetLastError();
}

falx_status check_input(const falx_ctx* c, const uint8_t* d, uint64_t len) {
    if (!c) return FALX_E_INVALID;
    if (len >= (1ULL << 32)) return FALX_E_TOO_LARGE;
    if (len > 0 && (!d || ((uintptr_t)d & 15u))) return FALX_E_INVALID;
    return FALX_OK;
}

// The spec's part of the Params of a launch over d_data[0, len) (no scratch).
void spec_params(const falx_spec* spec, const uint8_t* d, uint64_t len, Params& P) {
    memset(&P, 0, sizeof P);
    P.data = d;
    P.len = (uint32_t)len;
    P.n_tiles = (uint32_t)((len + TILE - 1) / TILE);
    for (int k = 0; k < 4; ++k) P.sepb[k] = 0x100;
    if (spec) {
        for (uint32_t i = 0; i < spec->n_structural; ++i) {
            uint8_t b = spec->structural[i];
            if (b != '\n') P.cand[b >> 5] |= 1u << (b & 31);
        }
        bool nl = false, sep = false;
        for (uint32_t i = 0; i < spec->n_structural; ++i) (spec->structural[i] == '\n' ? nl : sep) = true;
        if (nl || !sep) P.cand['\n' >> 5] |= 1u << ('\n' & 31);
        // Up to four candidate bytes besides '\n' are tested by compares
        // instead of the table; '\n' comes from the newline test.
        for (uint32_t b = 0; b < 256; ++b)
            if ((P.cand[b >> 5] >> (b & 31)) & 1u) {
                if (b == '\n') {
                    P.nl_cand = 1;
                    continue;
                }
                if (P.n_sepb < 4) P.sepb[P.n_sepb] = b;
                ++P.n_sepb;
            }
        P.quote = spec->quote;
        P.escape = spec->quote >= 0 ? spec->escape : -1;
        P.comment = spec->comment;
        P.n_cols = spec->n_columns;
        uint32_t ns = 0;
        for (uint32_t k = 0; k < spec->n_columns; ++k) {
            P.cols[k].index = spec->columns[k].index;
            P.cols[k].type = spec->columns[k].type;
            P.cols[k].str_slot = spec->columns[k].type == FALX_COL_STR ? ns++ : 0;
        }
        P.n_str = ns;
    }
}

// Common Params for a launch over d_data[0, len), scratch in the context.
falx_status prepare(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len, Params& P) {
    spec_params(spec, d, len, P);
    const uint32_t ns = kKindStates[spec ? dialect_kind(spec) : 0];
    falx_status st = ensure_lb(c, P.n_tiles, P.n_str, ns);
    if (st != FALX_OK) return st;
    LbLayout L = lb_layout(P.n_tiles, P.n_str, ns);
    P.str_len = reinterpret_cast<uint32_t*>(static_cast<char*>(c->d_lb) + L.str_len);
    P.str_base = reinterpret_cast<const uint32_t*>(static_cast<char*>(c->d_lb) + L.str_base);
    P.result = c->d_result;
    return FALX_OK;
}

hipError_t fill_u64(hipStream_t s, unsigned long long* p, uint64_t n, unsigned long long v) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + WG - 1) / WG;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)blocks), dim3(WG), 0, s, p, n, v);
    return hipGetLastError();
}

// A device-to-host read on the context's stream (never the null stream,
// whose implicit synchronisation would break graph captures running on
// other streams of the process).
hipError_t read_back(falx_ctx* c, void* h, const void* d, size_t n) {
    hipError_t e = hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, c->stream);
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}

falx_status begin(falx_ctx* c) {
    tl_hip_error = 0;
    (void)hipGetLastError();  // launches below report their own errors only
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(fill_u64(c->stream, c->d_result, R_SLOTS, 0ULL));
    c->timed = false;
    return FALX_OK;
}

// Events bracket the call's main kernel(s) only (fills and the final
// reductions excluded); falx_ctx_last_kernel_ms reports that span.
hipError_t mark_start(falx_ctx* c) { return hipEventRecord(c->ev0, c->stream); }

hipError_t mark_end(falx_ctx* c) {
    c->timed = true;
    return hipEventRecord(c->ev1, c->stream);
}

// Bring the result slots home and synchronise.
falx_status finish(falx_ctx* c) {
    if (!c->timed) {
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        HIP_TRY(hipEventRecord(c->ev1, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(c->h_result, c->d_result, R_SLOTS * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return FALX_OK;
}

// Two-level fixed-order reductions into c->d_red + RED_OUT.
hipError_t reduce_u64_2level(falx_ctx* c, const unsigned long long* parts, uint32_t n, uint32_t width,
                             size_t out_off = 0) {
    unsigned long long* tmp = static_cast<unsigned long long*>(c->d_red);
    unsigned long long* out = reinterpret_cast<unsigned long long*>(static_cast<char*>(c->d_red) + RED_OUT + out_off);
    uint32_t blocks = n < RED_BLOCKS * 4 ? 1u : RED_BLOCKS;
    hipLaunchKernelGGL(k_reduce_u64, dim3(blocks), dim3(WG), 0, c->stream, parts, n, width, blocks == 1 ? out : tmp);
    if (blocks > 1) hipLaunchKernelGGL(k_reduce_u64, dim3(1), dim3(WG), 0, c->stream, tmp, blocks, width, out);
    return hipGetLastError();
}

hipError_t reduce_stats_2level(falx_ctx* c, const StatPart* parts, uint32_t n, uint32_t n_cols) {
    StatPart* tmp = static_cast<StatPart*>(c->d_red);
    StatPart* out = reinterpret_cast<StatPart*>(static_cast<char*>(c->d_red) + RED_OUT);
    uint32_t blocks = n < RED_BLOCKS * 4 ? 1u : RED_BLOCKS;
    hipLaunchKernelGGL(k_reduce_stats, dim3(n_cols, blocks), dim3(WG), 0, c->stream, parts, n, n_cols,
                       blocks == 1 ? out : tmp);
    if (blocks > 1)
        hipLaunchKernelGGL(k_reduce_stats, dim3(n_cols, 1), dim3(WG), 0, c->stream, tmp, blocks, n_cols, out);
    return hipGetLastError();
}

}  // namespace

extern "C" {

falx_status falx_ctx_create(int device, falx_ctx** out) {
    if (!out) return FALX_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n) return FALX_E_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return FALX_E_NO_DEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return FALX_E_NO_DEVICE;
    falx_ctx* c = new (std::nothrow) falx_ctx();
    if (!c) return FALX_E_DEVICE;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_count, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&c->d_result, R_SLOTS * 8) != hipSuccess ||
        hipMalloc(&c->d_red, RED_BYTES) != hipSuccess ||
        hipHostMalloc(&c->h_result, R_SLOTS * 8, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->h_red), H_RED_BYTES, hipHostMallocDefault) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        falx_ctx_destroy(c);
        return FALX_E_DEVICE;
    }
    *out = c;
    return FALX_OK;
}

void falx_ctx_destroy(falx_ctx* c) {
    if (!c) return;
    // Teardown is best effort: errors here have no caller to report to.
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->d_lb) (void)hipFree(c->d_lb);
    if (c->d_part) (void)hipFree(c->d_part);
    if (c->d_red) (void)hipFree(c->d_red);
    if (c->d_result) (void)hipFree(c->d_result);
    if (c->h_result) (void)hipHostFree(c->h_result);
    if (c->h_red) (void)hipHostFree(c->h_red);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev_count) (void)hipEventDestroy(c->ev_count);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void* falx_ctx_stream(falx_ctx* c) { return c ? (void*)c->stream : nullptr; }

double falx_ctx_last_kernel_ms(falx_ctx* c) { return c ? (double)c->last_ms : 0.0; }

falx_status falx_ctx_last_error(const falx_ctx* c, int32_t* hip_error) {
    if (!c) return FALX_E_INVALID;
    if (hip_error) *hip_error = tl_hip_error;
    return FALX_OK;
}

falx_status falx_device_alloc(falx_ctx* c, uint64_t bytes, void** d_out) {
    if (!c || !d_out) return FALX_E_INVALID;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMalloc(d_out, bytes ? bytes : 1));
    return FALX_OK;
}

falx_status falx_device_free(falx_ctx* c, void* p) {
    if (!c) return FALX_E_INVALID;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (p) HIP_TRY(hipFree(p));
    return FALX_OK;
}

falx_status falx_host_alloc_pinned(falx_ctx* c, uint64_t bytes, void** h_out) {
    if (!c || !h_out) return FALX_E_INVALID;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipHostMalloc(h_out, bytes ? bytes : 1, hipHostMallocDefault));
    return FALX_OK;
}

falx_status falx_host_free_pinned(falx_ctx* c, void* p) {
    if (!c) return FALX_E_INVALID;
    if (p) HIP_TRY(hipHostFree(p));
    return FALX_OK;
}

falx_status falx_upload(falx_ctx* c, void* d_dst, const void* h_src, uint64_t bytes) {
    if (!c) return FALX_E_INVALID;
    if (!bytes) return FALX_OK;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FALX_OK;
}

falx_status falx_download(falx_ctx* c, void* h_dst, const void* d_src, uint64_t bytes) {
    if (!c) return FALX_E_INVALID;
    if (!bytes) return FALX_OK;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FALX_OK;
}

falx_status falx_spec_validate(const falx_spec* s) {
    if (!s) return FALX_E_INVALID;
    if (s->n_structural == 0 || s->n_structural > FALX_MAX_STRUCTURAL) return FALX_E_INVALID;
    if (s->quote < -1 || s->quote > 255 || s->escape < -1 || s->escape > 255 || s->comment < -1 || s->comment > 255)
        return FALX_E_INVALID;
    if (s->quote >= 0 && s->escape == s->quote) return FALX_E_INVALID;       // formats.rs:237-240
    if (s->comment == '\n' || s->comment == '\r') return FALX_E_INVALID;      // spec.rs comment rules
    if (s->comment >= 0 && s->comment == s->quote) return FALX_E_INVALID;
    if (s->n_columns > FALX_MAX_COLUMNS) return FALX_E_INVALID;               // codegen.rs:198-205
    for (uint32_t k = 0; k < s->n_columns; ++k)
        if (s->columns[k].type > FALX_COL_BYTES) return FALX_E_INVALID;
    return FALX_OK;
}

falx_status falx_index_structurals(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len,
                                   uint32_t* d_out, uint64_t out_cap, uint64_t* n_out) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if ((st = falx_spec_validate(spec)) != FALX_OK || !n_out) return st != FALX_OK ? st : FALX_E_INVALID;
    *n_out = 0;
    if (len == 0) return FALX_OK;
    Params P;
    if ((st = begin(c)) != FALX_OK) return st;
    if ((st = prepare(c, spec, d, len, P)) != FALX_OK) return st;
    P.out = d_out;
    P.out_cap = d_out ? out_cap : 0;
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, dialect_kind(spec)));
    HIP_TRY(launch_scan<M_INDEX>(c, P, dialect_kind(spec)));
    HIP_TRY(mark_end(c));
    if ((st = finish(c)) != FALX_OK) return st;
    *n_out = c->h_result[R_N_STRUCT];
    return *n_out > P.out_cap ? FALX_E_CAPACITY : FALX_OK;
}

falx_status falx_parse(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len, uint32_t* d_seps,
                       uint64_t seps_cap, uint64_t* d_ends, uint64_t ends_cap, uint64_t* n_seps, uint64_t* n_ends) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if ((st = falx_spec_validate(spec)) != FALX_OK) return st;
    if (!n_seps || !n_ends) return FALX_E_INVALID;
    *n_seps = *n_ends = 0;
    if (len == 0) return FALX_OK;
    Params P;
    if ((st = begin(c)) != FALX_OK) return st;
    if ((st = prepare(c, spec, d, len, P)) != FALX_OK) return st;
    P.out = d_seps;
    P.out_cap = d_seps ? seps_cap : 0;
    P.ends = reinterpret_cast<unsigned long long*>(d_ends);
    P.ends_cap = d_ends ? ends_cap : 0;
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, dialect_kind(spec)));
    HIP_TRY(launch_scan<M_TAPE>(c, P, dialect_kind(spec)));
    HIP_TRY(mark_end(c));
    if ((st = finish(c)) != FALX_OK) return st;
    *n_seps = c->h_result[R_N_SEPS];
    *n_ends = c->h_result[R_N_TERMS];
    return (*n_seps > P.out_cap || *n_ends > P.ends_cap) ? FALX_E_CAPACITY : FALX_OK;
}

static falx_status reduce_stats(falx_ctx* c, const falx_spec* spec, uint32_t n_tiles, falx_csv_stats* out);

static falx_status columns_impl(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len,
                                falx_columns* out, falx_csv_stats* stats) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if ((st = falx_spec_validate(spec)) != FALX_OK) return st;
    if (!out) return FALX_E_INVALID;
    for (uint32_t k = 0; k < spec->n_columns; ++k) {
        const falx_column_buf& b = out->cols[k];
        if (!b.values || (!b.valid && out->rows_cap)) return FALX_E_INVALID;
        if (spec->columns[k].type == FALX_COL_STR && !b.data && b.data_cap) return FALX_E_INVALID;
    }
    out->rows = 0;
    if (stats) memset(stats, 0, sizeof *stats);
    Params P;
    if ((st = begin(c)) != FALX_OK) return st;
    // Validity words are OR-ed by the kernel: clear them first; string
    // offsets start at 0 even for empty input.
    for (uint32_t k = 0; k < spec->n_columns; ++k) {
        HIP_TRY(fill_u64(c->stream, reinterpret_cast<unsigned long long*>(out->cols[k].valid),
                         (out->rows_cap + 63) / 64, 0ULL));
        out->cols[k].data_len = 0;
    }
    if (len == 0) {
        for (uint32_t k = 0; k < spec->n_columns; ++k)
            if (spec->columns[k].type == FALX_COL_STR)
                HIP_TRY(hipMemsetAsync(out->cols[k].values, 0, 4, c->stream));
        return finish(c);
    }
    if ((st = prepare(c, spec, d, len, P)) != FALX_OK) return st;
    P.rows_cap = out->rows_cap;
    for (uint32_t k = 0; k < spec->n_columns; ++k) {
        P.cols[k].values = out->cols[k].values;
        P.cols[k].valid = reinterpret_cast<unsigned long long*>(out->cols[k].valid);
        P.cols[k].data = out->cols[k].data;
        P.cols[k].data_cap = out->cols[k].data_cap;
    }
    if (stats) {
        uint32_t nc = spec->n_columns ? spec->n_columns : 1;
        size_t part_bytes = (size_t)P.n_tiles * nc * sizeof(StatPart) + (size_t)P.n_tiles * 8;
        if ((st = ensure_part(c, part_bytes)) != FALX_OK) return st;
        P.partials = c->d_part;
    }
    const int kind = dialect_kind(spec);
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, kind));
    if (numeric_fast(spec, kind, 0, spec->n_columns)) {
        if (stats) HIP_TRY((launch_fast<M_NUM_FAST_STATS, M_COLUMNS>(c, P, kind)));
        else HIP_TRY((launch_fast<M_NUM_FAST, M_COLUMNS>(c, P, kind)));
    } else {
        if (stats) HIP_TRY((launch_fast<M_COLUMNS_FAST_STATS, M_COLUMNS>(c, P, kind)));
        else HIP_TRY((launch_fast<M_COLUMNS_FAST, M_COLUMNS>(c, P, kind)));
    }
    // Statistics of the columns neither fused pass covers come from extra
    // passes over their groups.
    if (stats) {
        Params Q = P;
        for (uint32_t g = 0; g < spec->n_columns; g += STAT_GROUP) {
            if (!stats_group_needed(spec, g)) continue;
            Q.col_group = g;
            HIP_TRY(launch_stats_group(c, spec, Q, kind));
        }
    }
    // String columns: arena bases from the per-tile byte counts, then the
    // placement pass.
    if (P.n_str) {
        hipLaunchKernelGGL(k_slot_scan_wide, dim3((P.n_tiles + SLOT_CHUNK - 1) / SLOT_CHUNK, P.n_str), dim3(SLOT_WG), 0,
                           c->stream, P.str_len, P.n_tiles, P.n_str,
                           const_cast<uint32_t*>(P.str_base), P.result);
        HIP_TRY(hipGetLastError());
        HIP_TRY(launch_scan<M_STRINGS>(c, P, dialect_kind(spec)));
    }
    HIP_TRY(mark_end(c));
    if (stats) {
        if ((st = reduce_stats(c, spec, P.n_tiles, stats)) != FALX_OK) return st;
    } else if ((st = finish(c)) != FALX_OK) {
        return st;
    }
    out->rows = c->h_result[R_N_ROWS];
    bool cap = out->rows > out->rows_cap;
    for (uint32_t k = 0; k < spec->n_columns; ++k) {
        if (spec->columns[k].type != FALX_COL_STR) continue;
        uint64_t dl = c->h_result[R_STR_LEN + P.cols[k].str_slot];
        out->cols[k].data_len = dl;
        if (dl > out->cols[k].data_cap) cap = true;
        if (dl > 0x7FFFFFFFULL) return FALX_E_TOO_LARGE;
    }
    return cap ? FALX_E_CAPACITY : FALX_OK;
}

// Reduce the per-tile StatParts + record counts in c->d_part, finish the
// call and fill `out`.
static falx_status reduce_stats(falx_ctx* c, const falx_spec* spec, uint32_t n_tiles, falx_csv_stats* out) {
    const StatPart* parts = static_cast<const StatPart*>(c->d_part);
    if (spec->n_columns) HIP_TRY(reduce_stats_2level(c, parts, n_tiles, spec->n_columns));
    // record counts: a u64 per tile after the StatParts, summed the same way
    HIP_TRY(reduce_u64_2level(c, reinterpret_cast<const unsigned long long*>(parts + (size_t)n_tiles * spec->n_columns),
                              n_tiles, 1u, REC_OFF));
    // The reduced StatParts and record count come home with the result
    // slots, under the one synchronisation of finish().
    const char* red = static_cast<const char*>(c->d_red) + RED_OUT;
    const size_t sp_bytes = sizeof(StatPart) * spec->n_columns;
    static_assert(sizeof(StatPart) * FALX_MAX_COLUMNS + 8 <= H_RED_BYTES, "reduction outputs fit their pinned home");
    if (sp_bytes) HIP_TRY(hipMemcpyAsync(c->h_red, red, sp_bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_red + sizeof(StatPart) * FALX_MAX_COLUMNS, red + REC_OFF, 8, hipMemcpyDeviceToHost,
                           c->stream));
    falx_status st = finish(c);
    if (st != FALX_OK) return st;
    StatPart host[FALX_MAX_COLUMNS];
    if (sp_bytes) memcpy(host, c->h_red, sp_bytes);
    unsigned long long records = 0;
    memcpy(&records, c->h_red + sizeof(StatPart) * FALX_MAX_COLUMNS, 8);
    out->records = records;
    for (uint32_t k = 0; k < spec->n_columns; ++k) {
        out->cols[k].values = host[k].values;
        out->cols[k].checksum = host[k].checksum;
        out->cols[k].bytes = host[k].bytes;
        out->cols[k].fixed6_sum = host[k].fixed6_sum;
        out->cols[k].fixed6_cells = host[k].fixed6_cells;
        out->cols[k].sum = host[k].sum;
    }
    return FALX_OK;
}

falx_status falx_parse_columns(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len,
                               falx_columns* out) {
    return columns_impl(c, spec, d, len, out, nullptr);
}

falx_status falx_parse_columns_with_stats(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len,
                                          falx_columns* out, falx_csv_stats* stats) {
    if (!stats) return FALX_E_INVALID;
    return columns_impl(c, spec, d, len, out, stats);
}

falx_status falx_parse_csv_stats(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len,
                                 falx_csv_stats* out) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if ((st = falx_spec_validate(spec)) != FALX_OK) return st;
    if (!out) return FALX_E_INVALID;
    memset(out, 0, sizeof *out);
    if (len == 0) return FALX_OK;
    Params P;
    if ((st = begin(c)) != FALX_OK) return st;
    if ((st = prepare(c, spec, d, len, P)) != FALX_OK) return st;
    uint32_t nc = spec->n_columns ? spec->n_columns : 1;
    size_t part_bytes = (size_t)P.n_tiles * nc * sizeof(StatPart) + (size_t)P.n_tiles * 8;
    if ((st = ensure_part(c, part_bytes)) != FALX_OK) return st;
    P.partials = c->d_part;
    P.n_cols = spec->n_columns;
    uint32_t groups = spec->n_columns ? (spec->n_columns + STAT_GROUP - 1) / STAT_GROUP : 1;
    const int kind = dialect_kind(spec);
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, kind));
    for (uint32_t g = 0; g < groups; ++g) {
        P.col_group = g * STAT_GROUP;
        HIP_TRY(launch_stats_group(c, spec, P, kind));
    }
    HIP_TRY(mark_end(c));
    return reduce_stats(c, spec, P.n_tiles, out);
}

falx_status falx_parse_field_bytes(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len,
                                   falx_field_byte_stats* out) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if ((st = falx_spec_validate(spec)) != FALX_OK) return st;
    if (!out) return FALX_E_INVALID;
    out->fields = out->bytes = 0;
    if (len == 0) return FALX_OK;
    Params P;
    if ((st = begin(c)) != FALX_OK) return st;
    if ((st = prepare(c, spec, d, len, P)) != FALX_OK) return st;
    if ((st = ensure_part(c, (size_t)P.n_tiles * 16)) != FALX_OK) return st;
    P.partials = c->d_part;
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, dialect_kind(spec)));
    HIP_TRY(launch_scan<M_FIELDBYTES>(c, P, dialect_kind(spec)));
    HIP_TRY(mark_end(c));
    HIP_TRY(reduce_u64_2level(c, static_cast<const unsigned long long*>(c->d_part), P.n_tiles, 2u));
    const char* red = static_cast<const char*>(c->d_red) + RED_OUT;
    if ((st = finish(c)) != FALX_OK) return st;
    unsigned long long h[2];
    HIP_TRY(read_back(c, h, red, 16));
    out->fields = h[0];
    out->bytes = h[1];
    return FALX_OK;
}

static falx_spec fixed_spec(const char* structural, int quote, int escape) {
    falx_spec s;
    memset(&s, 0, sizeof s);
    s.n_structural = (uint32_t)strlen(structural);
    memcpy(s.structural, structural, s.n_structural);
    s.quote = quote;
    s.escape = escape;
    s.comment = -1;
    return s;
}

falx_status falx_parse_ndjson_lines(falx_ctx* c, const uint8_t* d, uint64_t len, uint64_t* records) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if (!records) return FALX_E_INVALID;
    *records = 0;
    if (len == 0) return FALX_OK;
    falx_spec spec = fixed_spec("\n", '"', '\\');
    Params P;
    if ((st = begin(c)) != FALX_OK) return st;
    if ((st = prepare(c, &spec, d, len, P)) != FALX_OK) return st;
    // The tile summaries already count the terminators: the scan's grand
    // total is the answer (no emit pass).
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, dialect_kind(&spec)));
    HIP_TRY(mark_end(c));
    if ((st = finish(c)) != FALX_OK) return st;
    uint8_t last = 0;
    HIP_TRY(read_back(c, &last, d + len - 1, 1));
    *records = c->h_result[R_N_TERMS] + (last != '\n' ? 1 : 0);  // ndjson_has_trailing_record (ndjson.rs:853-855)
    return FALX_OK;
}

falx_status falx_parse_ndjson_id_score(falx_ctx* c, const uint8_t* d, uint64_t len,
                                       falx_ndjson_id_score_stats* out) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if (!out) return FALX_E_INVALID;
    memset(out, 0, sizeof *out);
    if (len == 0) return FALX_OK;
    if ((st = begin(c)) != FALX_OK) return st;
    uint32_t n_tiles = (uint32_t)((len + TILE - 1) / TILE);
    // Per tile: the tile's partial, then (second half) its far record's.
    if ((st = ensure_part(c, (size_t)2 * n_tiles * sizeof(NdPart) + (size_t)n_tiles * sizeof(uint2))) != FALX_OK)
        return st;
    NdPart* parts = static_cast<NdPart*>(c->d_part);
    uint2* far = reinterpret_cast<uint2*>(parts + 2 * (size_t)n_tiles);
    HIP_TRY(mark_start(c));
    hipLaunchKernelGGL(k_ndjson_id_score, dim3((n_tiles + ND_WAVES - 1) / ND_WAVES), dim3(ND_WG), 0, c->stream, d,
                       (uint32_t)len, n_tiles,
                       parts, far);
    hipLaunchKernelGGL(k_ndjson_far, dim3((n_tiles + 63) / 64), dim3(64), 0, c->stream, d, n_tiles,
                       static_cast<const uint2*>(far), parts + n_tiles);
    HIP_TRY(mark_end(c));
    {
        NdPart* tmp = static_cast<NdPart*>(c->d_red);
        NdPart* fin = reinterpret_cast<NdPart*>(static_cast<char*>(c->d_red) + RED_OUT);
        const uint32_t n_parts = 2 * n_tiles;
        uint32_t blocks = n_parts < RED_BLOCKS * 4 ? 1u : RED_BLOCKS;
        hipLaunchKernelGGL(k_reduce_nd, dim3(blocks), dim3(WG), 0, c->stream, static_cast<const NdPart*>(parts),
                           n_parts, blocks == 1 ? fin : tmp);
        if (blocks > 1) hipLaunchKernelGGL(k_reduce_nd, dim3(1), dim3(WG), 0, c->stream, tmp, blocks, fin);
        HIP_TRY(hipGetLastError());
    }
    const NdPart* red = reinterpret_cast<const NdPart*>(static_cast<const char*>(c->d_red) + RED_OUT);
    if ((st = finish(c)) != FALX_OK) return st;
    NdPart h;
    HIP_TRY(read_back(c, &h, red, sizeof h));
    out->records = h.records;
    out->invalid = h.invalid;
    out->id_sum = h.id_sum;
    out->score_fixed6_sum = h.fx_sum;
    out->score_fixed6_cells = h.fx_cells;
    out->score_sum = h.score_sum;
    return out->invalid ? FALX_E_SCHEMA : FALX_OK;
}

falx_status falx_parse_logfmt_pairs(falx_ctx* c, const uint8_t* d, uint64_t len, falx_logfmt_stats* out) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if (!out) return FALX_E_INVALID;
    memset(out, 0, sizeof *out);
    if (len == 0) return FALX_OK;
    falx_spec spec = fixed_spec(" =\n", '"', '\\');
    Params P;
    if ((st = begin(c)) != FALX_OK) return st;
    if ((st = prepare(c, &spec, d, len, P)) != FALX_OK) return st;
    if ((st = ensure_part(c, (size_t)P.n_tiles * 32)) != FALX_OK) return st;
    P.partials = c->d_part;
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, dialect_kind(&spec)));
    HIP_TRY(launch_scan<M_LOGFMT>(c, P, dialect_kind(&spec)));
    HIP_TRY(mark_end(c));
    HIP_TRY(reduce_u64_2level(c, static_cast<const unsigned long long*>(c->d_part), P.n_tiles, 4u));
    const char* red = static_cast<const char*>(c->d_red) + RED_OUT;
    if ((st = finish(c)) != FALX_OK) return st;
    unsigned long long h[4];
    HIP_TRY(read_back(c, h, red, 32));
    out->pairs = h[0];
    out->key_bytes = h[1];
    out->value_bytes = h[2];
    out->checksum = h[3];
    return FALX_OK;
}

}  // extern "C"

namespace {

// Scratch of the bracket matcher over n_tiles tiles of NT index entries.
struct NestLayout {
    size_t tsum, tmin, tmis, rcount, res, bstat, base, lv[NT_LEVELS], total;
    uint32_t lv_n[NT_LEVELS];
};

NestLayout nest_layout(uint32_t n_tiles) {
    NestLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    L.tsum = take(4 * (size_t)n_tiles);
    L.tmin = take(4 * (size_t)n_tiles);
    L.tmis = take(4 * (size_t)n_tiles);
    L.rcount = take(4 * (size_t)n_tiles);
    L.res = take(4 * (size_t)NT * n_tiles);
    L.bstat = take(4 * (size_t)NT_BLOCKS * n_tiles);
    L.base = take(8 * (size_t)n_tiles);
    uint32_t m = n_tiles;
    for (int k = 0; k < NT_LEVELS; ++k) {
        L.lv_n[k] = m;
        L.lv[k] = take(8 * (size_t)m);
        m = (m + 63) / 64;
    }
    L.total = o;
    return L;
}

}  // namespace

extern "C" {

falx_status falx_parse_nested(falx_ctx* c, const uint8_t* d, uint64_t len, uint64_t* d_tape, uint64_t tape_cap,
                              uint64_t* n_tape, int32_t* error, uint64_t* error_pos) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if (!n_tape || !error || !error_pos) return FALX_E_INVALID;
    *n_tape = 0;
    *error = FALX_NEST_OK;
    *error_pos = 0;
    if (len == 0) return FALX_OK;
    // The JSON dialect of json.rs: `"` strings with backslash escapes,
    // structural bytes , : [ ] { } (formats.rs:164-173).
    falx_spec spec = fixed_spec(",:[]{}", '"', '\\');
    const int kind = dialect_kind(&spec);
    Params P;
    if ((st = begin(c)) != FALX_OK) return st;
    if ((st = prepare(c, &spec, d, len, P)) != FALX_OK) return st;
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, kind));
    // The entry count sizes the matcher's grid and scratch: it is copied
    // home behind the prefix while the index pass (which needs only the
    // tape's capacity) runs.
    HIP_TRY(hipMemcpyAsync(c->h_result, c->d_result, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(c->ev_count, c->stream));
    unsigned long long* tape = reinterpret_cast<unsigned long long*>(d_tape);
    P.ends = tape;
    P.ends_cap = d_tape ? tape_cap : 0;
    if (d_tape) HIP_TRY(launch_scan<M_INDEX>(c, P, kind));
    HIP_TRY(hipEventSynchronize(c->ev_count));
    const uint64_t n = c->h_result[R_N_STRUCT];
    if (n == 0 || n > P.ends_cap) {
        HIP_TRY(mark_end(c));
        if ((st = finish(c)) != FALX_OK) return st;
        *n_tape = n;
        return n == 0 ? FALX_OK : FALX_E_CAPACITY;
    }
    const uint32_t n_tiles = (uint32_t)((n + NT - 1) / NT);
    const NestLayout L = nest_layout(n_tiles);
    if ((st = ensure_part(c, L.total)) != FALX_OK) return st;
    char* base = static_cast<char*>(c->d_part);
    NestScratch S;
    S.tsum = reinterpret_cast<int32_t*>(base + L.tsum);
    S.tmin = reinterpret_cast<int32_t*>(base + L.tmin);
    S.tmis = reinterpret_cast<uint32_t*>(base + L.tmis);
    S.rcount = reinterpret_cast<uint32_t*>(base + L.rcount);
    S.res = reinterpret_cast<uint32_t*>(base + L.res);
    S.bstat = reinterpret_cast<uint32_t*>(base + L.bstat);
    S.base = reinterpret_cast<long long*>(base + L.base);
    for (int k = 0; k < NT_LEVELS; ++k) {
        S.lv[k] = reinterpret_cast<long long*>(base + L.lv[k]);
        S.lv_n[k] = L.lv_n[k];
    }
    S.result = c->d_result;
    hipLaunchKernelGGL(k_nest_local, dim3(n_tiles), dim3(WG), 0, c->stream, tape, (uint32_t)n, S);
    hipLaunchKernelGGL(k_nest_prefix, dim3(1), dim3(NT_PREFIX_WG), 0, c->stream, n_tiles, S);
    for (int k = 1; k < NT_LEVELS; ++k)
        hipLaunchKernelGGL(k_nest_tree, dim3((L.lv_n[k] + WG - 1) / WG), dim3(WG), 0, c->stream,
                           static_cast<const long long*>(S.lv[k - 1]), L.lv_n[k - 1], S.lv[k], L.lv_n[k]);
    hipLaunchKernelGGL(k_nest_far, dim3(n_tiles), dim3(NT_FAR_WG), 0, c->stream, d, tape, (uint32_t)n, S);
    HIP_TRY(hipGetLastError());
    HIP_TRY(mark_end(c));
    if ((st = finish(c)) != FALX_OK) return st;
    const uint64_t cut = std::min(c->h_result[R_NT_NEG], c->h_result[R_NT_MIS]);
    const uint64_t open = c->h_result[R_NT_OPEN];
    const long long depth = (long long)c->h_result[R_NT_DEPTH];
    *n_tape = n;
    uint64_t at = UINT64_MAX;
    if (cut < n) {
        // json.rs:733-752: the tape stops before the first unmatched close.
        *n_tape = cut;
        *error = FALX_NEST_UNMATCHED_CLOSE;
        at = cut;
        if (cut > 0) {
            const uint32_t blocks = (uint32_t)std::min<uint64_t>((cut + WG - 1) / WG, 65536);
            hipLaunchKernelGGL(k_nest_cut, dim3(blocks), dim3(WG), 0, c->stream, tape, (uint32_t)cut);
            HIP_TRY(hipGetLastError());
        }
    } else if (depth > 0 && open > 0) {
        // json.rs:651-653: the innermost open still open at the end.
        *error = FALX_NEST_UNCLOSED_OPEN;
        at = open - 1;
    }
    if (at != UINT64_MAX) {
        unsigned long long v = 0;
        HIP_TRY(read_back(c, &v, tape + at, 8));
        *error_pos = (uint32_t)v;
    }
    return FALX_OK;
}

// ---------------------------------------------------------------------------
// Streaming ingestion (StreamParser, csv_geo.rs:688-807): the unfinished
// record of the previous feed is kept at the start of a device buffer, the
// next chunk is appended behind it, and the tape of that buffer is built
// with the plain pipeline. The bytes after a structural terminator start in
// the dialect's entry state (outside quotes, no pending escape, at a line
// start), so re-indexing the kept record from its first byte gives the
// whole document's records exactly. Records completed by the feed are
// returned; the bytes after the last terminator move to the other buffer.
// ---------------------------------------------------------------------------
struct falx_stream {
    falx_ctx* ctx = nullptr;
    falx_spec spec{};
    uint8_t* buf[2] = {nullptr, nullptr};
    uint64_t cap = 0;
    int cur = 0;          // the buffer holding the unfinished record
    uint64_t tail = 0;    // its length, at buf[cur][0, tail)
    bool finished = false;
};

falx_status falx_stream_create(falx_ctx* c, const falx_spec* spec, uint64_t buf_cap, falx_stream** out) {
    if (!c || !out) return FALX_E_INVALID;
    *out = nullptr;
    falx_status st = falx_spec_validate(spec);
    if (st != FALX_OK) return st;
    if (buf_cap == 0) return FALX_E_INVALID;
    if (buf_cap >= (1ULL << 32)) return FALX_E_TOO_LARGE;
    falx_stream* s = new (std::nothrow) falx_stream();
    if (!s) return FALX_E_DEVICE;
    s->ctx = c;
    s->spec = *spec;
    s->cap = buf_cap;
    if (hipSetDevice(c->device) != hipSuccess || hipMalloc(&s->buf[0], buf_cap) != hipSuccess ||
        hipMalloc(&s->buf[1], buf_cap) != hipSuccess) {
        falx_stream_destroy(s);
        return FALX_E_DEVICE;
    }
    *out = s;
    return FALX_OK;
}

falx_status falx_stream_reset(falx_stream* s) {
    if (!s) return FALX_E_INVALID;
    s->tail = 0;
    s->cur = 0;
    s->finished = false;
    return FALX_OK;
}

void falx_stream_destroy(falx_stream* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (uint8_t* b : s->buf)
        if (b) (void)hipFree(b);
    delete s;
}

namespace {

// The tape of buf[cur][0, total) into the caller's arrays; `last` true for
// the final call (the unterminated last record is emitted too).
falx_status stream_emit(falx_stream* s, uint64_t total, bool last, uint32_t* d_seps, uint64_t seps_cap,
                        uint64_t* d_ends, uint64_t ends_cap, falx_stream_out* out) {
    falx_ctx* c = s->ctx;
    const uint8_t* b = s->buf[s->cur];
    uint64_t n_seps = 0, n_ends = 0;
    falx_status st = falx_parse(c, &s->spec, b, total, d_seps, seps_cap, d_ends, ends_cap, &n_seps, &n_ends);
    out->d_buf = b;
    if (st == FALX_E_CAPACITY) {
        out->n_seps = n_seps;
        out->n_ends = n_ends;
        return st;
    }
    if (st != FALX_OK) return st;
    uint64_t keep_from = total;   // bytes from here on wait for the next feed
    if (!last) {
        if (n_ends == 0) {
            keep_from = 0;
            n_seps = 0;
        } else {
            unsigned long long w = 0;
            HIP_TRY(read_back(c, &w, d_ends + n_ends - 1, 8));
            keep_from = (w & 0xFFFFFFFFull) + 1;
            n_seps = w >> 32;
        }
    }
    out->n_seps = n_seps;
    out->n_ends = n_ends;
    out->data_end = keep_from;
    const uint64_t tail = total - keep_from;
    if (tail && keep_from) {
        const int nxt = 1 - s->cur;
        HIP_TRY(hipMemcpyAsync(s->buf[nxt], b + keep_from, tail, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        s->cur = nxt;
    }
    s->tail = tail;
    return FALX_OK;
}

}  // namespace

falx_status falx_stream_feed(falx_stream* s, const uint8_t* d_chunk, uint64_t len, uint32_t* d_seps,
                             uint64_t seps_cap, uint64_t* d_ends, uint64_t ends_cap, falx_stream_out* out) {
    if (!s || !out || s->finished || (len && !d_chunk)) return FALX_E_INVALID;
    memset(out, 0, sizeof *out);
    falx_ctx* c = s->ctx;
    tl_hip_error = 0;
    const uint64_t total = s->tail + len;
    if (total > s->cap) {
        out->need_bytes = total;
        return FALX_E_CAPACITY;
    }
    HIP_TRY(hipSetDevice(c->device));
    if (len) HIP_TRY(hipMemcpyAsync(s->buf[s->cur] + s->tail, d_chunk, len, hipMemcpyDeviceToDevice, c->stream));
    // A capacity failure leaves the stream as it was (the chunk is copied
    // again by the retry).
    return stream_emit(s, total, false, d_seps, seps_cap, d_ends, ends_cap, out);
}

falx_status falx_stream_finish(falx_stream* s, uint32_t* d_seps, uint64_t seps_cap, uint64_t* d_ends,
                               uint64_t ends_cap, falx_stream_out* out) {
    if (!s || !out || s->finished) return FALX_E_INVALID;
    memset(out, 0, sizeof *out);
    tl_hip_error = 0;
    out->d_buf = s->buf[s->cur];
    if (s->tail == 0) {
        s->finished = true;
        return FALX_OK;
    }
    falx_status st = stream_emit(s, s->tail, true, d_seps, seps_cap, d_ends, ends_cap, out);
    if (st == FALX_OK) s->finished = true;
    return st;
}

// ---------------------------------------------------------------------------
// BGZF (bgzf.rs:61-290): block boundaries on the host, inflation on the
// device (inflate.h).
// ---------------------------------------------------------------------------
static_assert(sizeof(falx_bgzf_block) == sizeof(InfBlock), "block table layout");

falx_status falx_bgzf_scan(const uint8_t* h, uint64_t len, falx_bgzf_block* blocks, uint64_t cap,
                           uint64_t* n_blocks, uint64_t* total_out) {
    if ((!h && len) || !n_blocks || !total_out) return FALX_E_INVALID;
    *n_blocks = *total_out = 0;
    constexpr uint64_t HEADER = 12, TRAILER = 8;
    uint64_t pos = 0, n = 0, out = 0;
    while (pos < len) {
        // gzip member with FEXTRA (bgzf.rs:64-86)
        if (pos + HEADER > len || h[pos] != 0x1f || h[pos + 1] != 0x8b || h[pos + 2] != 8 || !(h[pos + 3] & 0x04))
            return FALX_E_INVALID;
        const uint64_t xlen = (uint64_t)h[pos + 10] | ((uint64_t)h[pos + 11] << 8);
        const uint64_t extra_end = pos + HEADER + xlen;
        if (extra_end > len) return FALX_E_INVALID;
        // the BC subfield holds BSIZE = block length - 1 (bgzf.rs:40-57)
        uint64_t bsize = UINT64_MAX;
        for (uint64_t i = pos + HEADER; i + 4 <= extra_end;) {
            const uint64_t slen = (uint64_t)h[i + 2] | ((uint64_t)h[i + 3] << 8);
            if (h[i] == 'B' && h[i + 1] == 'C' && slen == 2 && i + 6 <= extra_end) {
                bsize = (uint64_t)h[i + 4] | ((uint64_t)h[i + 5] << 8);
                break;
            }
            i += 4 + slen;
        }
        if (bsize == UINT64_MAX) return FALX_E_INVALID;
        const uint64_t block_len = bsize + 1;
        if (block_len <= HEADER + xlen + TRAILER || pos + block_len > len) return FALX_E_INVALID;
        const uint64_t end = pos + block_len;
        const uint32_t isize = (uint32_t)h[end - 4] | ((uint32_t)h[end - 3] << 8) | ((uint32_t)h[end - 2] << 16) |
                               ((uint32_t)h[end - 1] << 24);
        if (isize > 0) {   // the empty EOF block is dropped
            if (n < cap) {
                blocks[n].payload = extra_end;
                blocks[n].payload_len = (uint32_t)(end - TRAILER - extra_end);
                blocks[n].isize = isize;
                blocks[n].out = out;
            }
            ++n;
            out += isize;
        }
        pos = end;
    }
    *n_blocks = n;
    *total_out = out;
    return n > cap ? FALX_E_CAPACITY : FALX_OK;
}

falx_status falx_bgzf_inflate(falx_ctx* c, const uint8_t* d_comp, uint64_t comp_len, const falx_bgzf_block* blocks,
                              uint64_t n_blocks, uint8_t* d_out, uint64_t out_cap, uint64_t* first_bad) {
    if (!c || (n_blocks && (!blocks || !d_comp || !d_out))) return FALX_E_INVALID;
    if (first_bad) *first_bad = UINT64_MAX;
    if (n_blocks == 0) return FALX_OK;
    if (n_blocks >= (1ull << 31)) return FALX_E_TOO_LARGE;
    // Every payload inside the compressed bytes, every slot inside the output.
    for (uint64_t k = 0; k < n_blocks; ++k) {
        const falx_bgzf_block& b = blocks[k];
        if (b.payload > comp_len || b.payload_len > comp_len - b.payload || b.out > out_cap ||
            b.isize > out_cap - b.out)
            return FALX_E_INVALID;
    }
    falx_status st;
    if ((st = begin(c)) != FALX_OK) return st;
    const size_t table = n_blocks * sizeof(falx_bgzf_block);
    if ((st = ensure_part(c, table)) != FALX_OK) return st;
    HIP_TRY(hipMemcpyAsync(c->d_part, blocks, table, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(fill_u64(c->stream, c->d_result + R_INF_BAD, 1, ~0ull));
    HIP_TRY(mark_start(c));
    hipLaunchKernelGGL(k_inflate, dim3((uint32_t)n_blocks), dim3(INF_WG), 0, c->stream, d_comp,
                       static_cast<const InfBlock*>(c->d_part), d_out, c->d_result + R_INF_BAD);
    HIP_TRY(hipGetLastError());
    HIP_TRY(mark_end(c));
    if ((st = finish(c)) != FALX_OK) return st;
    const unsigned long long bad = c->h_result[R_INF_BAD];
    if (bad != ~0ull) {
        if (first_bad) *first_bad = bad;
        return FALX_E_INVALID;   // bgzf.rs: InvalidData
    }
    return FALX_OK;
}

// ---------------------------------------------------------------------------
// Index without allocation: the scratch of falx_index_structurals (tile
// summaries, group summaries, prefixes, result slots) in one file-scope
// device array, so the call makes no allocation at all.
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kStaticMaxTiles = FALX_STATIC_MAX_LEN / TILE;
constexpr LbLayout kStaticLayout = lb_layout(kStaticMaxTiles, 0, NS_MAX);
constexpr size_t kStaticResult = align_up(kStaticLayout.total, 256);
constexpr size_t kStaticBytes = kStaticResult + 8 * R_SLOTS;
static_assert(FALX_STATIC_MAX_LEN % TILE == 0, "whole tiles");

// Each device holds its own copy until the process ends (about 2 MiB).
__device__ __align__(256) uint8_t scratch[kStaticBytes];

// Device address of `scratch` per device (the name is a host stand-in), set
// once per device by hipGetSymbolAddress with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

falx_status static_scratch(int device, char** out) {
    if (device < 0 || device >= kMaxDevices) return FALX_E_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return FALX_E_DEVICE;
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<char*>(g_scratch_ptr[device]);
    return FALX_OK;
}

}  // namespace

falx_status falx_index_structurals_static(falx_ctx* c, const falx_spec* spec, const uint8_t* d, uint64_t len,
                                          uint32_t* d_out, uint64_t out_cap, uint64_t* n_out) {
    falx_status st = check_input(c, d, len);
    if (st != FALX_OK) return st;
    if ((st = falx_spec_validate(spec)) != FALX_OK) return st;
    if (!n_out) return FALX_E_INVALID;
    *n_out = 0;
    if (len > FALX_STATIC_MAX_LEN) return FALX_E_TOO_LARGE;
    if (len == 0) return FALX_OK;
    tl_hip_error = 0;
    (void)hipGetLastError();
    HIP_TRY(hipSetDevice(c->device));
    char* base = nullptr;
    if ((st = static_scratch(c->device, &base)) != FALX_OK) return st;
    unsigned long long* result = reinterpret_cast<unsigned long long*>(base + kStaticResult);
    Params P;
    spec_params(spec, d, len, P);
    P.n_cols = P.n_str = 0;   // the index reads no columns
    P.result = result;
    P.out = d_out;
    P.out_cap = d_out ? out_cap : 0;
    // Result slots start at zero: cleared by a fill kernel, not a memset.
    HIP_TRY(fill_u64(c->stream, result, R_SLOTS, 0ULL));
    HIP_TRY(mark_start(c));
    HIP_TRY(run_prefix(c, P, dialect_kind(spec), base));
    HIP_TRY(launch_scan<M_INDEX>(c, P, dialect_kind(spec)));
    HIP_TRY(mark_end(c));
    HIP_TRY(hipMemcpyAsync(c->h_result, result, R_SLOTS * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    *n_out = c->h_result[R_N_STRUCT];
    return *n_out > P.out_cap ? FALX_E_CAPACITY : FALX_OK;
}

// ---------------------------------------------------------------------------
// Arrow export (Arrow C Device Data Interface, zero-copy over the column).
// ---------------------------------------------------------------------------
namespace {

// What the export allocates on the host for one array/schema pair; the
// release callbacks free it (the column buffers are the caller's).
struct ArrowPriv {
    const void* buffers[3];
    const void* child_buffers[2];
    ArrowArray child;
    ArrowArray* child_ptr;
    char name[24];
};

struct ArrowSchemaPriv {
    ArrowSchema child;
    ArrowSchema* child_ptr;
    char name[24];
};

void arrow_release_array(ArrowArray* a) {
    if (!a || !a->release) return;
    if (a->private_data) delete static_cast<ArrowPriv*>(a->private_data);
    a->release = nullptr;
    a->private_data = nullptr;
}

// Children are released with their parent (they share its allocation).
void arrow_release_child_array(ArrowArray* a) {
    if (a) a->release = nullptr;
}

void arrow_release_schema(ArrowSchema* s) {
    if (!s || !s->release) return;
    if (s->private_data) delete static_cast<ArrowSchemaPriv*>(s->private_data);
    s->release = nullptr;
    s->private_data = nullptr;
}

void arrow_release_child_schema(ArrowSchema* s) {
    if (s) s->release = nullptr;
}

}  // namespace

falx_status falx_export_arrow(const falx_spec* spec, const falx_columns* cols, uint32_t k, int32_t device_type,
                              int64_t device_id, ArrowDeviceArray* out, ArrowSchema* schema) {
    falx_status st = falx_spec_validate(spec);
    if (st != FALX_OK) return st;
    if (!cols || !out || !schema || k >= spec->n_columns) return FALX_E_INVALID;
    if (cols->rows > cols->rows_cap || cols->rows > (uint64_t)INT32_MAX) return FALX_E_INVALID;
    const falx_column_buf& b = cols->cols[k];
    const uint32_t type = spec->columns[k].type;
    ArrowPriv* ap = new (std::nothrow) ArrowPriv();
    ArrowSchemaPriv* sp = new (std::nothrow) ArrowSchemaPriv();
    if (!ap || !sp) {
        delete ap;
        delete sp;
        return FALX_E_DEVICE;
    }
    const int64_t rows = (int64_t)cols->rows;
    snprintf(sp->name, sizeof sp->name, "c%u", spec->columns[k].index);
    memset(out, 0, sizeof *out);
    memset(schema, 0, sizeof *schema);
    ArrowArray& a = out->array;
    a.length = rows;
    a.null_count = -1;
    a.offset = 0;
    a.buffers = ap->buffers;
    a.release = arrow_release_array;
    a.private_data = ap;
    schema->name = sp->name;
    schema->flags = ARROW_FLAG_NULLABLE;
    schema->release = arrow_release_schema;
    schema->private_data = sp;
    ap->buffers[0] = b.valid;
    if (type == FALX_COL_I64 || type == FALX_COL_F64) {
        schema->format = type == FALX_COL_I64 ? "l" : "g";
        a.n_buffers = 2;
        ap->buffers[1] = b.values;
    } else if (type == FALX_COL_STR) {
        schema->format = "z";
        a.n_buffers = 3;
        ap->buffers[1] = b.values;
        ap->buffers[2] = b.data;
    } else {
        // (start, end) pairs: a fixed-size list of two uint32 per cell.
        schema->format = "+w:2";
        a.n_buffers = 1;
        a.n_children = 1;
        ap->child_buffers[0] = nullptr;
        ap->child_buffers[1] = b.values;
        ap->child.length = 2 * rows;
        ap->child.null_count = 0;
        ap->child.n_buffers = 2;
        ap->child.buffers = ap->child_buffers;
        ap->child.release = arrow_release_child_array;
        ap->child_ptr = &ap->child;
        a.children = &ap->child_ptr;
        schema->n_children = 1;
        sp->child.format = "I";
        snprintf(sp->name + 12, 12, "span");
        sp->child.name = sp->name + 12;
        sp->child.release = arrow_release_child_schema;
        sp->child_ptr = &sp->child;
        schema->children = &sp->child_ptr;
    }
    out->device_id = device_id;
    out->device_type = device_type;
    out->sync_event = nullptr;
    return FALX_OK;
}

}  // extern "C"

#include "batch.hip"
