// batch.hip — batched execution of many documents (included by falx_gpu.hip).
//
// A batch runs each stage of the plain pipeline once for all of its
// documents: persistent workgroups drain a device queue of (document,
// chunk-of-tiles) units, so documents of uneven sizes and mixed specs share
// every launch. The stage sequence reads its work from a descriptor blob at a
// fixed device address (uploaded in one copy per batch), so it is captured
// once per set of stages in a hipGraph and replayed for every later batch
// with that set. Results are bit-identical to the plain entry points: the
// tile kernels are the same and the reductions keep their fixed orders.

namespace {

constexpr uint32_t RES_SLOTS_OFF = 0;                                   // u64[R_SLOTS]
constexpr uint32_t RES_STATS_OFF = 8 * R_SLOTS;                         // StatPart[32]
constexpr uint32_t RES_REC_OFF = RES_STATS_OFF + 48 * FALX_MAX_COLUMNS; // u64 records
constexpr uint32_t RES_U64_OFF = RES_REC_OFF + 8;                       // u64[4]
constexpr uint32_t RES_ND_OFF = RES_U64_OFF + 32;                       // NdPart
constexpr uint32_t RES_STRIDE = (RES_ND_OFF + (uint32_t)sizeof(NdPart) + 255) & ~255u;
static_assert(sizeof(StatPart) == 48, "StatPart layout");

// Size classes for the default chunk (tiles per queue unit): class k holds
// documents of [4^k, 4^(k+1)) tiles.
constexpr uint32_t N_SIZE_CLASS = 8;

uint32_t size_class(uint32_t n_tiles) {
    uint32_t k = 0;
    while (k + 1 < N_SIZE_CLASS && n_tiles >= (1u << (2 * (k + 1)))) ++k;
    return k;
}

bool op_uses_spec(uint32_t op) { return op >= FALX_OP_INDEX && op <= FALX_OP_FIELD_BYTES; }

template <int MODE>
hipError_t launch_tiles_batch(int kind, dim3 g, hipStream_t s, const BatchHdr* H, uint32_t tab) {
    return launch_tile_kernel(kind, F_BATCH, MODE, g, s, nullptr, H, tab);
}

hipError_t launch_emit(uint32_t e, int kind, dim3 g, hipStream_t s, const BatchHdr* H, uint32_t tab) {
    switch (e) {
    case E_INDEX: return launch_tiles_batch<M_INDEX>(kind, g, s, H, tab);
    case E_COLUMNS: return launch_tiles_batch<M_COLUMNS>(kind, g, s, H, tab);
    case E_STATS: return launch_tiles_batch<M_STATS>(kind, g, s, H, tab);
    case E_STRINGS: return launch_tiles_batch<M_STRINGS>(kind, g, s, H, tab);
    case E_LOGFMT: return launch_tiles_batch<M_LOGFMT>(kind, g, s, H, tab);
    default: return launch_tiles_batch<M_FIELDBYTES>(kind, g, s, H, tab);
    }
}

// Host image of the descriptor blob, built per batch.
struct BlobBuilder {
    std::vector<Params> params;
    std::vector<BPrefix> pre;
    std::vector<BRed> red;
    std::vector<BNd> nd;
    std::vector<BFill> fill;
    struct Tab {
        std::vector<uint32_t> req, chunk, start;
        uint32_t total = 0;
        void add(uint32_t r, uint32_t units, uint32_t c) {
            if (!units) return;
            req.push_back(r);
            chunk.push_back(c);
            start.push_back(total);
            total += units;
        }
    } tab[N_TAB];
};

// One document of the batch being assembled or in flight.
struct Doc {
    falx_request r;
    falx_spec spec;          // copy (logfmt/ndjson: the fixed dialect)
    uint32_t n_tiles = 0, kind = 0, n_str = 0;
    uint32_t pidx = 0;       // its main Params
    uint32_t res = 0;        // index of its results record
    int32_t early = -1;      // status decided on the host (empty input, bad argument)
};

}  // namespace

struct falx_batch {
    falx_ctx* c = nullptr;
    char* d_scratch = nullptr;
    size_t scratch_cap = 0;
    size_t scratch_budget = 0;   // what admission may use (the caller's figure)
    char* d_blob = nullptr;
    size_t blob_cap = 0;
    char* h_blob = nullptr;
    char* h_res = nullptr;
    size_t h_res_cap = 0;
    size_t res_off = 0, res_bytes = 0;
    std::vector<Doc> docs;
    bool inflight = false;
    hipEvent_t done = nullptr;
    falx_batch_tuning tune{};
    std::vector<std::pair<uint64_t, hipGraphExec_t>> graphs;
    uint32_t n_cu = 256;
    float last_ms = 0.f;
    hipEvent_t t0 = nullptr;
};

namespace {

void drop_graphs(falx_batch* b) {
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
}

uint64_t scratch_of(const falx_request* r, uint32_t* n_tiles_out = nullptr) {
    if (!r || r->len == 0) return 0;
    const uint32_t n_tiles = (uint32_t)((r->len + TILE - 1) / TILE);
    if (n_tiles_out) *n_tiles_out = n_tiles;
    if (r->op == FALX_OP_NDJSON_ID_SCORE)
        return align_up((size_t)2 * n_tiles * sizeof(NdPart), 256) + align_up((size_t)n_tiles * sizeof(uint2), 256) +
               align_up(RED_BLOCKS * sizeof(NdPart), 256);
    falx_spec fixed;
    const falx_spec* spec = r->spec;
    if (r->op == FALX_OP_LOGFMT_PAIRS) {
        fixed = fixed_spec(" =\n", '"', '\\');
        spec = &fixed;
    }
    if (!spec) return 0;
    uint32_t n_str = 0;
    for (uint32_t k = 0; k < spec->n_columns && k < FALX_MAX_COLUMNS; ++k) n_str += spec->columns[k].type == FALX_COL_STR;
    const LbLayout L = lb_layout(n_tiles, n_str, kKindStates[dialect_kind(spec)]);
    const uint32_t nc = spec->n_columns ? spec->n_columns : 1;
    size_t part = 0, tmp = 0;
    switch (r->op) {
    case FALX_OP_COLUMNS_STATS:
    case FALX_OP_CSV_STATS:
        part = (size_t)n_tiles * nc * sizeof(StatPart) + (size_t)n_tiles * 8;
        tmp = (size_t)RED_BLOCKS * nc * sizeof(StatPart) + (size_t)RED_BLOCKS * 8;
        break;
    case FALX_OP_FIELD_BYTES: part = (size_t)n_tiles * 16; tmp = RED_BLOCKS * 16; break;
    case FALX_OP_LOGFMT_PAIRS: part = (size_t)n_tiles * 32; tmp = RED_BLOCKS * 32; break;
    default: break;
    }
    return align_up(L.total, 256) + align_up(part, 256) + align_up(tmp, 256) + 256;
}

// Appends the reduction job of `out_off` bytes into the results record.
uint32_t add_red(BlobBuilder& B, uint32_t kind, const void* parts, uint32_t n, uint32_t width, void* tmp, void* out) {
    BRed J;
    J.parts = parts;
    J.tmp = tmp;
    J.out = out;
    J.kind = kind;
    J.n = n;
    J.width = width;
    J.blocks = n < RED_BLOCKS * 4 ? 1u : RED_BLOCKS;
    const uint32_t id = (uint32_t)B.red.size();
    B.red.push_back(J);
    const uint32_t per = kind == RJ_STATS ? width : 1u;
    B.tab[T_RED1].add(id, per * J.blocks, 1);
    if (J.blocks > 1) B.tab[T_RED2].add(id, per, 1);
    return id;
}

falx_status ensure_dev(falx_batch* b, char*& d, size_t& cap, size_t need, bool blob) {
    if (need <= cap) return FALX_OK;
    if (d) HIP_TRY(hipFree(d));
    d = nullptr;
    cap = 0;
    const size_t n = align_up(need + need / 4, 1 << 20);
    HIP_TRY(hipMalloc(&d, n));
    cap = n;
    if (blob) {
        if (b->h_blob) HIP_TRY(hipHostFree(b->h_blob));
        b->h_blob = nullptr;
        HIP_TRY(hipHostMalloc(&b->h_blob, n));
        drop_graphs(b);   // the header moved
    }
    return FALX_OK;
}

}  // namespace

extern "C" {

uint64_t falx_request_scratch_bytes(const falx_request* r) { return scratch_of(r); }

falx_status falx_batch_create(falx_ctx* c, uint64_t scratch_bytes, falx_batch** out) {
    if (!c || !out) return FALX_E_INVALID;
    *out = nullptr;
    falx_batch* b = new (std::nothrow) falx_batch;
    if (!b) return FALX_E_INVALID;
    b->c = c;
    tl_hip_error = 0;
    hipDeviceProp_t prop;
    if (hipSetDevice(c->device) != hipSuccess || hipGetDeviceProperties(&prop, c->device) != hipSuccess ||
        hipEventCreate(&b->done) != hipSuccess || hipEventCreate(&b->t0) != hipSuccess) {
        delete b;
        return FALX_E_DEVICE;
    }
    b->n_cu = (uint32_t)prop.multiProcessorCount;
    b->tune.tile_workgroups = 2 * b->n_cu;
    b->tune.unit_workgroups = b->n_cu;
    for (uint32_t k = 0; k < N_SIZE_CLASS; ++k) b->tune.chunk[k] = k < 3 ? 1u : (k < 5 ? 2u : 4u);
    b->tune.use_graph = 1;
    b->scratch_budget = scratch_bytes;
    if (scratch_bytes) {
        falx_status st = ensure_dev(b, b->d_scratch, b->scratch_cap, scratch_bytes, false);
        if (st != FALX_OK) {
            falx_batch_destroy(b);
            return st;
        }
    }
    *out = b;
    return FALX_OK;
}

void falx_batch_destroy(falx_batch* b) {
    if (!b) return;
    if (b->inflight) (void)hipEventSynchronize(b->done);
    drop_graphs(b);
    if (b->d_scratch) (void)hipFree(b->d_scratch);
    if (b->d_blob) (void)hipFree(b->d_blob);
    if (b->h_blob) (void)hipHostFree(b->h_blob);
    if (b->h_res) (void)hipHostFree(b->h_res);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->t0) (void)hipEventDestroy(b->t0);
    delete b;
}

falx_status falx_batch_get_tuning(const falx_batch* b, falx_batch_tuning* t) {
    if (!b || !t) return FALX_E_INVALID;
    *t = b->tune;
    return FALX_OK;
}

falx_status falx_batch_set_tuning(falx_batch* b, const falx_batch_tuning* t) {
    if (!b || !t || b->inflight) return FALX_E_INVALID;
    if (t->tile_workgroups == 0 || t->unit_workgroups == 0 || t->tile_workgroups > 65535 ||
        t->unit_workgroups > 65535)
        return FALX_E_INVALID;
    for (uint32_t k = 0; k < N_SIZE_CLASS; ++k)
        if (t->chunk[k] == 0 || t->chunk[k] > 64) return FALX_E_INVALID;
    const bool regrid = t->tile_workgroups != b->tune.tile_workgroups || t->unit_workgroups != b->tune.unit_workgroups;
    b->tune = *t;
    if (regrid) drop_graphs(b);
    return FALX_OK;
}

falx_status falx_batch_submit(falx_batch* b, const falx_request* reqs, uint32_t n) {
    if (!b || (n && !reqs)) return FALX_E_INVALID;
    if (b->inflight) return FALX_E_BUSY;
    falx_ctx* c = b->c;
    tl_hip_error = 0;
    (void)hipGetLastError();
    HIP_TRY(hipSetDevice(c->device));
    b->docs.assign(n, Doc{});
    BlobBuilder B;
    size_t scr = 0;
    // ---- documents: validation, scratch, Params, tables ----------------------
    for (uint32_t i = 0; i < n; ++i) {
        Doc& D = b->docs[i];
        D.r = reqs[i];
        D.res = i;
        const falx_request& r = D.r;
        falx_status st = check_input(c, r.d_data, r.len);
        if (st == FALX_OK && (r.op < FALX_OP_INDEX || r.op > FALX_OP_NDJSON_ID_SCORE)) st = FALX_E_INVALID;
        if (st == FALX_OK && op_uses_spec(r.op)) st = r.spec ? falx_spec_validate(r.spec) : FALX_E_INVALID;
        if (st == FALX_OK && r.op == FALX_OP_INDEX && r.d_index == nullptr && r.index_cap) st = FALX_E_INVALID;
        const bool cols = r.op == FALX_OP_COLUMNS || r.op == FALX_OP_COLUMNS_STATS;
        if (st == FALX_OK && cols) {
            if (!r.columns) st = FALX_E_INVALID;
            else
                for (uint32_t k = 0; k < r.spec->n_columns; ++k) {
                    const falx_column_buf& cb = r.columns->cols[k];
                    if (!cb.values || (!cb.valid && r.columns->rows_cap)) st = FALX_E_INVALID;
                    if (r.spec->columns[k].type == FALX_COL_STR && !cb.data && cb.data_cap) st = FALX_E_INVALID;
                }
        }
        if (st != FALX_OK) {
            D.early = st;
            continue;
        }
        if (op_uses_spec(r.op)) D.spec = *r.spec;
        else if (r.op == FALX_OP_LOGFMT_PAIRS) D.spec = fixed_spec(" =\n", '"', '\\');
        else D.spec = fixed_spec("\n", '"', '\\');
        uint32_t n_tiles = 0;
        const size_t need = scratch_of(&r, &n_tiles);
        if (r.len && scr + need > b->scratch_budget) {
            D.early = FALX_E_BUSY;   // not admitted: the batch's scratch budget is spent
            continue;
        }
        if (cols) {
            // Validity words are OR-ed by the kernels: cleared first, as the
            // plain entry point does.
            for (uint32_t k = 0; k < D.spec.n_columns; ++k) {
                const uint64_t words = (r.columns->rows_cap + 63) / 64;
                if (!words) continue;
                B.fill.push_back(BFill{reinterpret_cast<unsigned long long*>(r.columns->cols[k].valid), words, 0ULL});
                B.tab[T_FILL].add((uint32_t)B.fill.size() - 1, (uint32_t)((words + FILL_UNIT - 1) / FILL_UNIT), 1);
            }
        }
        if (r.len == 0) {
            D.early = FALX_OK;
            continue;
        }
        char* base = b->d_scratch + scr;
        scr += need;
        D.n_tiles = n_tiles;
        const uint32_t chunk = r.chunk ? std::min<uint32_t>(r.chunk, 64u) : b->tune.chunk[size_class(n_tiles)];
        // Offset of its results record; pointers into it are patched to blob
        // addresses once the layout is known.
        const size_t res_rel = (size_t)i * RES_STRIDE;
        if (r.op == FALX_OP_NDJSON_ID_SCORE) {
            // Per tile its partial, then (second half) its far record's.
            NdPart* parts = reinterpret_cast<NdPart*>(base);
            const size_t far_off = align_up((size_t)2 * n_tiles * sizeof(NdPart), 256);
            uint2* far = reinterpret_cast<uint2*>(base + far_off);
            NdPart* tmp = reinterpret_cast<NdPart*>(base + far_off + align_up((size_t)n_tiles * sizeof(uint2), 256));
            B.nd.push_back(BNd{r.d_data, (uint32_t)r.len, n_tiles, parts, far});
            // Units of whole blocks of the NDJSON kernel (ND_WAVES tiles).
            const uint32_t nd_chunk = (chunk + ND_WAVES - 1) / ND_WAVES * ND_WAVES;
            B.tab[T_NDJ].add((uint32_t)B.nd.size() - 1, (n_tiles + nd_chunk - 1) / nd_chunk, nd_chunk);
            add_red(B, RJ_ND, parts, 2 * n_tiles, 1, tmp, reinterpret_cast<void*>(res_rel + RES_ND_OFF));
            continue;
        }
        // CSV-family and logfmt documents: the tile-prefix pipeline.
        Params P;
        spec_params(&D.spec, r.d_data, r.len, P);
        D.kind = (uint32_t)dialect_kind(&D.spec);
        D.n_str = P.n_str;
        const uint32_t ns = kKindStates[D.kind];
        const LbLayout L = lb_layout(n_tiles, P.n_str, ns);
        P.tsum = reinterpret_cast<uint32_t*>(base + L.tsum);
        P.tpre = reinterpret_cast<const uint32_t*>(base + L.tpre);
        P.ts_stride = summary_stride(ns);
        P.str_len = reinterpret_cast<uint32_t*>(base + L.str_len);
        P.str_base = reinterpret_cast<const uint32_t*>(base + L.str_base);
        P.result = reinterpret_cast<unsigned long long*>(res_rel + RES_SLOTS_OFF);   // patched
        char* part_base = base + align_up(L.total, 256);
        const uint32_t nc = D.spec.n_columns ? D.spec.n_columns : 1;
        size_t part_bytes = 0;
        if (r.op == FALX_OP_COLUMNS_STATS || r.op == FALX_OP_CSV_STATS) part_bytes = (size_t)n_tiles * nc * 48 + (size_t)n_tiles * 8;
        else if (r.op == FALX_OP_FIELD_BYTES) part_bytes = (size_t)n_tiles * 16;
        else if (r.op == FALX_OP_LOGFMT_PAIRS) part_bytes = (size_t)n_tiles * 32;
        char* tmp_base = part_base + align_up(part_bytes, 256);
        if (part_bytes) P.partials = part_base;
        if (r.op == FALX_OP_INDEX) {
            P.out = r.d_index;
            P.out_cap = r.d_index ? r.index_cap : 0;
        }
        if (cols) {
            P.rows_cap = r.columns->rows_cap;
            for (uint32_t k = 0; k < D.spec.n_columns; ++k) {
                P.cols[k].values = r.columns->cols[k].values;
                P.cols[k].valid = reinterpret_cast<unsigned long long*>(r.columns->cols[k].valid);
                P.cols[k].data = r.columns->cols[k].data;
                P.cols[k].data_cap = r.columns->cols[k].data_cap;
            }
            if (r.op == FALX_OP_COLUMNS) P.partials = nullptr;
        }
        D.pidx = (uint32_t)B.params.size();
        B.params.push_back(P);
        // summary + prefix
        B.tab[T_SUM + D.kind].add(D.pidx, (n_tiles + chunk - 1) / chunk, chunk);
        const uint32_t groups = (n_tiles + SG - 1) / SG;
        BPrefix A;
        A.tsum = P.tsum;
        A.gsum = reinterpret_cast<uint32_t*>(base + L.gsum);
        A.gpre = reinterpret_cast<uint32_t*>(base + L.gpre);
        A.tpre = reinterpret_cast<uint32_t*>(base + L.tpre);
        A.result = P.result;   // patched
        A.stride = P.ts_stride;
        A.n_tiles = n_tiles;
        A.ns = ns;
        A.groups = groups;
        B.pre.push_back(A);
        const uint32_t pre_id = (uint32_t)B.pre.size() - 1;
        B.tab[T_GROUPS].add(pre_id, groups, 1);
        B.tab[T_PREQ].add(pre_id, 1, 1);
        const uint32_t units = (n_tiles + chunk - 1) / chunk;
        auto emit = [&](uint32_t e, uint32_t pidx) { B.tab[T_EMIT + D.kind * N_EMIT + e].add(pidx, units, chunk); };
        auto stats_pass = [&](uint32_t g) {
            Params Q = P;
            Q.col_group = g;
            B.params.push_back(Q);
            emit(E_STATS, (uint32_t)B.params.size() - 1);
        };
        switch (r.op) {
        case FALX_OP_INDEX: emit(E_INDEX, D.pidx); break;
        case FALX_OP_COLUMNS:
        case FALX_OP_COLUMNS_STATS:
            emit(E_COLUMNS, D.pidx);
            if (r.op == FALX_OP_COLUMNS_STATS)
                for (uint32_t g = 0; g < D.spec.n_columns; g += STAT_GROUP)
                    if (stats_group_needed(&D.spec, g)) stats_pass(g);
            if (P.n_str) {
                B.tab[T_SLOT].add(D.pidx, P.n_str, 1);
                emit(E_STRINGS, D.pidx);
            }
            break;
        case FALX_OP_CSV_STATS: {
            const uint32_t ng = D.spec.n_columns ? (D.spec.n_columns + STAT_GROUP - 1) / STAT_GROUP : 1;
            for (uint32_t g = 0; g < ng; ++g) stats_pass(g * STAT_GROUP);
            break;
        }
        case FALX_OP_FIELD_BYTES: emit(E_FIELDBYTES, D.pidx); break;
        case FALX_OP_LOGFMT_PAIRS: emit(E_LOGFMT, D.pidx); break;
        default: break;
        }
        // reductions into the results record
        if (r.op == FALX_OP_COLUMNS_STATS || r.op == FALX_OP_CSV_STATS) {
            if (D.spec.n_columns)
                add_red(B, RJ_STATS, part_base, n_tiles, D.spec.n_columns, tmp_base,
                        reinterpret_cast<void*>(res_rel + RES_STATS_OFF));
            add_red(B, RJ_U64, part_base + (size_t)n_tiles * D.spec.n_columns * 48, n_tiles, 1,
                    tmp_base + (size_t)RED_BLOCKS * nc * 48, reinterpret_cast<void*>(res_rel + RES_REC_OFF));
        } else if (r.op == FALX_OP_FIELD_BYTES) {
            add_red(B, RJ_U64, part_base, n_tiles, 2, tmp_base, reinterpret_cast<void*>(res_rel + RES_U64_OFF));
        } else if (r.op == FALX_OP_LOGFMT_PAIRS) {
            add_red(B, RJ_U64, part_base, n_tiles, 4, tmp_base, reinterpret_cast<void*>(res_rel + RES_U64_OFF));
        }
    }
    // ---- blob layout -----------------------------------------------------
    size_t o = align_up(sizeof(BatchHdr), 256);
    const size_t par_off = o;
    o = align_up(o + sizeof(Params) * B.params.size(), 256);
    const size_t pre_off = o;
    o = align_up(o + sizeof(BPrefix) * B.pre.size(), 256);
    const size_t red_off = o;
    o = align_up(o + sizeof(BRed) * B.red.size(), 256);
    const size_t nd_off = o;
    o = align_up(o + sizeof(BNd) * B.nd.size(), 256);
    const size_t fill_off = o;
    o = align_up(o + sizeof(BFill) * B.fill.size(), 256);
    size_t tab_off[N_TAB];
    for (uint32_t t = 0; t < N_TAB; ++t) {
        tab_off[t] = o;
        o = align_up(o + 4 * (3 * B.tab[t].req.size() + 1), 16);
    }
    o = align_up(o, 256);
    b->res_off = o;
    b->res_bytes = (size_t)n * RES_STRIDE;
    o += b->res_bytes;
    falx_status st = ensure_dev(b, b->d_blob, b->blob_cap, o, true);
    if (st != FALX_OK) return st;
    if (b->res_bytes > b->h_res_cap) {
        if (b->h_res) HIP_TRY(hipHostFree(b->h_res));
        b->h_res = nullptr;
        b->h_res_cap = 0;
        HIP_TRY(hipHostMalloc(&b->h_res, align_up(b->res_bytes, 1 << 16)));
        b->h_res_cap = align_up(b->res_bytes, 1 << 16);
    }
    char* h = b->h_blob;
    char* d = b->d_blob;
    memset(h, 0, o);   // results records start at zero
    auto dres = [&](const void* rel) { return static_cast<void*>(d + b->res_off + reinterpret_cast<uintptr_t>(rel)); };
    for (Params& P : B.params) P.result = static_cast<unsigned long long*>(dres(P.result));
    for (BPrefix& A : B.pre) A.result = static_cast<unsigned long long*>(dres(A.result));
    for (BRed& J : B.red) J.out = dres(J.out);
    BatchHdr H;
    memset(&H, 0, sizeof H);
    H.params = reinterpret_cast<const Params*>(d + par_off);
    H.pre = reinterpret_cast<const BPrefix*>(d + pre_off);
    H.red = reinterpret_cast<const BRed*>(d + red_off);
    H.nd = reinterpret_cast<const BNd*>(d + nd_off);
    H.fill = reinterpret_cast<const BFill*>(d + fill_off);
    uint64_t sig = 0;
    for (uint32_t t = 0; t < N_TAB; ++t) {
        const auto& T = B.tab[t];
        const size_t m = T.req.size();
        uint32_t* w = reinterpret_cast<uint32_t*>(h + tab_off[t]);
        memcpy(w, T.req.data(), 4 * m);
        memcpy(w + m, T.chunk.data(), 4 * m);
        memcpy(w + 2 * m, T.start.data(), 4 * m);
        w[3 * m] = T.total;
        const uint32_t* dw = reinterpret_cast<const uint32_t*>(d + tab_off[t]);
        H.tab[t] = SegT{(uint32_t)m, T.total, dw, dw + m, dw + 2 * m};
        if (T.total) sig |= 1ULL << t;
    }
    memcpy(h, &H, sizeof H);
    memcpy(h + par_off, B.params.data(), sizeof(Params) * B.params.size());
    memcpy(h + pre_off, B.pre.data(), sizeof(BPrefix) * B.pre.size());
    memcpy(h + red_off, B.red.data(), sizeof(BRed) * B.red.size());
    memcpy(h + nd_off, B.nd.data(), sizeof(BNd) * B.nd.size());
    memcpy(h + fill_off, B.fill.data(), sizeof(BFill) * B.fill.size());
    // ---- submission ------------------------------------------------------
    hipStream_t s = c->stream;
    HIP_TRY(hipEventRecord(b->t0, s));
    HIP_TRY(hipMemcpyAsync(d, h, o, hipMemcpyHostToDevice, s));
    // Empty string columns of empty documents: offsets[0] = 0, as the plain
    // entry point writes (4 zero bytes from the results area, which starts at 0).
    for (const Doc& D : b->docs)
        if (D.early == FALX_OK && D.r.len == 0 &&
            (D.r.op == FALX_OP_COLUMNS || D.r.op == FALX_OP_COLUMNS_STATS))
            for (uint32_t k = 0; k < D.spec.n_columns; ++k)
                if (D.spec.columns[k].type == FALX_COL_STR)
                    HIP_TRY(hipMemcpyAsync(D.r.columns->cols[k].values, h + b->res_off, 4, hipMemcpyHostToDevice, s));
    const BatchHdr* dH = reinterpret_cast<const BatchHdr*>(d);
    const dim3 tg(b->tune.tile_workgroups), ug(b->tune.unit_workgroups);
    auto launch_all = [&](hipStream_t ss) -> hipError_t {
        // The tile launches check their own errors; the first one found is
        // returned (the others are checked at the end).
        hipError_t first = hipSuccess;
        auto keep = [&](hipError_t e) {
            if (first == hipSuccess) first = e;
        };
        if (sig & (1ULL << T_FILL)) hipLaunchKernelGGL(k_fill_batch, ug, dim3(WG), 0, ss, dH);
        for (int k = 0; k < 6; ++k)
            if (sig & (1ULL << (T_SUM + k))) keep(launch_tiles_batch<M_SUMMARY>(k, tg, ss, dH, T_SUM + k));
        if (sig & (1ULL << T_GROUPS)) {
            hipLaunchKernelGGL(k_prefix_reduce_batch, ug, dim3(SG), 0, ss, dH);
            hipLaunchKernelGGL(k_prefix_groups_batch, ug, dim3(64), 0, ss, dH);
            hipLaunchKernelGGL(k_prefix_tiles_batch, ug, dim3(SG), 0, ss, dH);
        }
        const uint32_t order[] = {E_INDEX, E_COLUMNS, E_STATS, E_LOGFMT, E_FIELDBYTES};
        for (uint32_t e : order)
            for (int k = 0; k < 6; ++k) {
                const uint32_t t = T_EMIT + k * N_EMIT + e;
                if (sig & (1ULL << t)) keep(launch_emit(e, k, tg, ss, dH, t));
            }
        if (sig & (1ULL << T_SLOT)) {
            hipLaunchKernelGGL(k_slot_scan_batch, ug, dim3(SLOT_WG), 0, ss, dH);
            for (int k = 0; k < 6; ++k) {
                const uint32_t t = T_EMIT + k * N_EMIT + E_STRINGS;
                if (sig & (1ULL << t)) keep(launch_emit(E_STRINGS, k, tg, ss, dH, t));
            }
        }
        if (sig & (1ULL << T_NDJ)) {
            hipLaunchKernelGGL(k_ndjson_batch, tg, dim3(ND_WG), 0, ss, dH);
            hipLaunchKernelGGL(k_ndjson_far_batch, ug, dim3(64), 0, ss, dH);
        }
        if (sig & (1ULL << T_RED1)) hipLaunchKernelGGL(k_reduce_batch, ug, dim3(WG), 0, ss, dH, 1u);
        if (sig & (1ULL << T_RED2)) hipLaunchKernelGGL(k_reduce_batch, ug, dim3(WG), 0, ss, dH, 2u);
        keep(hipGetLastError());
        return first;
    };
    static_assert(N_TAB <= 64, "signature bits");
    if (sig) {
        // Batches of at least use_graph documents run their launch sequence
        // as a graph, captured the first time the sequence is needed; smaller
        // ones (light load makes many, of varied stage sets) launch it as it
        // is, which costs less than instantiating a graph on their latency.
        if (b->tune.use_graph && n >= b->tune.use_graph) {
            hipGraphExec_t exec = nullptr;
            for (auto& g : b->graphs)
                if (g.first == sig) exec = g.second;
            if (!exec) {
                hipGraph_t graph = nullptr;
                HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                const hipError_t le = launch_all(s);
                const hipError_t ce = hipStreamEndCapture(s, &graph);
                if (le != hipSuccess || ce != hipSuccess) {
                    if (graph) (void)hipGraphDestroy(graph);
                    tl_hip_error = (int32_t)(le != hipSuccess ? le : ce);
                    return FALX_E_DEVICE;
                }
                const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_TRY(ie);
                b->graphs.emplace_back(sig, exec);
            }
            HIP_TRY(hipGraphLaunch(exec, s));
        } else {
            HIP_TRY(launch_all(s));
        }
    }
    if (b->res_bytes)
        HIP_TRY(hipMemcpyAsync(b->h_res, d + b->res_off, b->res_bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(b->done, s));
    b->inflight = true;
    return FALX_OK;
}

falx_status falx_batch_query(falx_batch* b) {
    if (!b) return FALX_E_INVALID;
    if (!b->inflight) return FALX_OK;
    const hipError_t e = hipEventQuery(b->done);
    if (e == hipSuccess) return FALX_OK;
    if (e == hipErrorNotReady) return FALX_E_BUSY;
    tl_hip_error = (int32_t)e;
    return FALX_E_DEVICE;
}

falx_status falx_batch_wait(falx_batch* b, falx_result* results) {
    if (!b) return FALX_E_INVALID;
    if (!b->inflight) return FALX_E_INVALID;
    HIP_TRY(hipEventSynchronize(b->done));
    b->inflight = false;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->t0, b->done) == hipSuccess) b->last_ms = ms;
    for (size_t i = 0; i < b->docs.size(); ++i) {
        Doc& D = b->docs[i];
        falx_result R;
        memset(&R, 0, sizeof R);
        const falx_request& r = D.r;
        const bool cols = r.op == FALX_OP_COLUMNS || r.op == FALX_OP_COLUMNS_STATS;
        if (cols && r.columns) {
            r.columns->rows = 0;
            for (uint32_t k = 0; k < FALX_MAX_COLUMNS; ++k) r.columns->cols[k].data_len = 0;
        }
        if (D.early >= 0) {
            R.status = D.early;
            if (results) results[i] = R;
            continue;
        }
        const char* rec = b->h_res + (size_t)D.res * RES_STRIDE;
        const unsigned long long* slots = reinterpret_cast<const unsigned long long*>(rec + RES_SLOTS_OFF);
        falx_status st = FALX_OK;
        switch (r.op) {
        case FALX_OP_INDEX:
            R.n_index = slots[R_N_STRUCT];
            if (st == FALX_OK && R.n_index > (r.d_index ? r.index_cap : 0)) st = FALX_E_CAPACITY;
            break;
        case FALX_OP_COLUMNS:
        case FALX_OP_COLUMNS_STATS:
        case FALX_OP_CSV_STATS: {
            if (r.op != FALX_OP_COLUMNS) {
                const StatPart* sp = reinterpret_cast<const StatPart*>(rec + RES_STATS_OFF);
                R.stats.records = *reinterpret_cast<const unsigned long long*>(rec + RES_REC_OFF);
                for (uint32_t k = 0; k < D.spec.n_columns; ++k) {
                    R.stats.cols[k].values = sp[k].values;
                    R.stats.cols[k].checksum = sp[k].checksum;
                    R.stats.cols[k].bytes = sp[k].bytes;
                    R.stats.cols[k].fixed6_sum = sp[k].fixed6_sum;
                    R.stats.cols[k].fixed6_cells = sp[k].fixed6_cells;
                    R.stats.cols[k].sum = sp[k].sum;
                }
            }
            if (r.op == FALX_OP_CSV_STATS) break;
            r.columns->rows = slots[R_N_ROWS];
            bool cap = r.columns->rows > r.columns->rows_cap;
            uint32_t slot = 0;
            for (uint32_t k = 0; k < D.spec.n_columns; ++k) {
                if (D.spec.columns[k].type != FALX_COL_STR) continue;
                const uint64_t dl = slots[R_STR_LEN + slot++];
                r.columns->cols[k].data_len = dl;
                if (dl > r.columns->cols[k].data_cap) cap = true;
                if (dl > 0x7FFFFFFFULL && st == FALX_OK) st = FALX_E_TOO_LARGE;
            }
            if (cap && st == FALX_OK) st = FALX_E_CAPACITY;
            break;
        }
        case FALX_OP_FIELD_BYTES: {
            const unsigned long long* u = reinterpret_cast<const unsigned long long*>(rec + RES_U64_OFF);
            R.field_bytes.fields = u[0];
            R.field_bytes.bytes = u[1];
            break;
        }
        case FALX_OP_LOGFMT_PAIRS: {
            const unsigned long long* u = reinterpret_cast<const unsigned long long*>(rec + RES_U64_OFF);
            R.logfmt.pairs = u[0];
            R.logfmt.key_bytes = u[1];
            R.logfmt.value_bytes = u[2];
            R.logfmt.checksum = u[3];
            break;
        }
        case FALX_OP_NDJSON_ID_SCORE: {
            NdPart p;
            memcpy(&p, rec + RES_ND_OFF, sizeof p);
            R.ndjson.records = p.records;
            R.ndjson.invalid = p.invalid;
            R.ndjson.id_sum = p.id_sum;
            R.ndjson.score_fixed6_sum = p.fx_sum;
            R.ndjson.score_fixed6_cells = p.fx_cells;
            R.ndjson.score_sum = p.score_sum;
            if (p.invalid) st = FALX_E_SCHEMA;
            break;
        }
        default: break;
        }
        R.status = st;
        if (results) results[i] = R;
    }
    return FALX_OK;
}

double falx_batch_last_ms(const falx_batch* b) { return b ? (double)b->last_ms : 0.0; }

}  // extern "C"
