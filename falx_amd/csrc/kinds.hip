// kinds.hip — the tile kernels of one dialect kind (FX_KIND, 0-5), compiled
// once per kind so the six kinds build in parallel. Each kind's launches go
// through launch_kind<FX_KIND> (declared in scan.hip).
#include "scan.hip"

#ifndef FX_KIND
#error "FX_KIND must be defined (0-5)"
#endif

namespace fx {

namespace {

constexpr bool K_HQ = FX_KIND >= 2;
constexpr bool K_ESC = FX_KIND == 3 || FX_KIND == 5;
constexpr bool K_CMT = FX_KIND == 1 || FX_KIND == 4 || FX_KIND == 5;
// Kinds with neither escape nor comment: the SWAR front and its numeric fast forms.
constexpr bool K_SWAR = !K_ESC && !K_CMT;

template <int MODE>
hipError_t plain(dim3 g, hipStream_t s, const Params* P) {
    hipLaunchKernelGGL((k_scan<K_HQ, K_ESC, K_CMT, MODE>), g, dim3(WG), 0, s, *P);
    return hipGetLastError();
}

// The SWAR kinds' dedicated summary and numeric fast passes (swar_tile.h).
template <int MODE>
hipError_t swar(dim3 g, hipStream_t s, const Params* P) {
    if constexpr (MODE == M_SUMMARY) hipLaunchKernelGGL((k_swar_summary<K_HQ>), g, dim3(WG), 0, s, *P);
    else hipLaunchKernelGGL((k_numfast<K_HQ, MODE>), g, dim3(WG), 0, s, *P);
    return hipGetLastError();
}

template <int MODE>
hipError_t redo(dim3 g, hipStream_t s, const Params* P) {
    hipLaunchKernelGGL((k_redo<K_HQ, K_ESC, K_CMT, MODE>), g, dim3(WG), 0, s, *P);
    return hipGetLastError();
}

template <int MODE>
hipError_t batch(dim3 g, hipStream_t s, const BatchHdr* H, uint32_t tab) {
    hipLaunchKernelGGL((k_tiles_batch<K_HQ, K_ESC, K_CMT, MODE>), g, dim3(WG), 0, s, H, tab);
    return hipGetLastError();
}

}  // namespace

template <>
hipError_t launch_kind<FX_KIND>(int form, int mode, dim3 g, hipStream_t s, const Params* P, const BatchHdr* H,
                                uint32_t tab) {
    // logfmt pairs run on the quote + backslash dialect only.
    constexpr bool LOGFMT = FX_KIND == 3;
    if (form == F_PLAIN) {
        switch (mode) {
        case M_SUMMARY:
            if constexpr (K_SWAR) return swar<M_SUMMARY>(g, s, P);
            else return plain<M_SUMMARY>(g, s, P);
        case M_INDEX: return plain<M_INDEX>(g, s, P);
        case M_TAPE: return plain<M_TAPE>(g, s, P);
        case M_COLUMNS_FAST: return plain<M_COLUMNS_FAST>(g, s, P);
        case M_COLUMNS_FAST_STATS: return plain<M_COLUMNS_FAST_STATS>(g, s, P);
        case M_STATS_FAST: return plain<M_STATS_FAST>(g, s, P);
        case M_NUM_FAST:
            if constexpr (K_SWAR) return swar<M_NUM_FAST>(g, s, P);
            break;
        case M_NUM_FAST_STATS:
            if constexpr (K_SWAR) return swar<M_NUM_FAST_STATS>(g, s, P);
            break;
        case M_NUM_STATS_FAST:
            if constexpr (K_SWAR) return swar<M_NUM_STATS_FAST>(g, s, P);
            break;
        case M_STRINGS: return plain<M_STRINGS>(g, s, P);
        case M_FIELDBYTES: return plain<M_FIELDBYTES>(g, s, P);
        case M_LOGFMT:
            if constexpr (LOGFMT) return plain<M_LOGFMT>(g, s, P);
            break;
        default: break;
        }
    } else if (form == F_REDO) {
        switch (mode) {
        case M_COLUMNS: return redo<M_COLUMNS>(g, s, P);
        case M_STATS: return redo<M_STATS>(g, s, P);
        default: break;
        }
    } else if (form == F_BATCH) {
        switch (mode) {
        case M_SUMMARY: return batch<M_SUMMARY>(g, s, H, tab);
        case M_INDEX: return batch<M_INDEX>(g, s, H, tab);
        case M_COLUMNS: return batch<M_COLUMNS>(g, s, H, tab);
        case M_STATS: return batch<M_STATS>(g, s, H, tab);
        case M_STRINGS: return batch<M_STRINGS>(g, s, H, tab);
        case M_FIELDBYTES: return batch<M_FIELDBYTES>(g, s, H, tab);
        case M_LOGFMT:
            if constexpr (LOGFMT) return batch<M_LOGFMT>(g, s, H, tab);
            break;
        default: break;
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace fx
