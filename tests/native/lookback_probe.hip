// lookback_probe.hip — a measurement program, not part of the library: what
// a single-pass decoupled look-back would cost the tile pipeline on one GPU.
//
// Each workgroup stands for one tile of 2 KiB per wave (NW waves: 8, 16 or
// 32 KiB). It issues a fixed amount of VALU work (A instructions per wave),
// publishes its aggregate (a transfer function over two entry states and
// four counters per state, the shape of a quote-only dialect's tile
// summary), looks back over its predecessors for an inclusive prefix, then
// issues the rest of its work (C per wave). Published words carry the call's
// epoch in their high half and are written and read with agent-scope
// relaxed atomics (global_store/global_load with sc1, no fences), so a
// reader accepts a word only when its tag is current. Wave 0 looks back
// over a window of 64, 128 or 256 predecessors (1, 2 or 4 per lane) and
// waits until the window holds an inclusive prefix; every spin is bounded
// and a shared abort flag ends the others. The result of every tile is
// checked against a serial scan on the host.
//
// Build and run (one GPU):
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -Xclang -target-feature \
//     -Xclang -packed-fp32-ops -o /tmp/lookback_probe tests/native/lookback_probe.hip
//   /tmp/lookback_probe 300 650
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <vector>
#include <cstdlib>

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP %s line %d\n", hipGetErrorString(e_), __LINE__); exit(1);} } while (0)

constexpr int NSTATE = 2, NCNT = 4;
constexpr int REC = 16;   // u64 words per tile record: agg[2][4] + fn + inc[4] + state (+pad)
constexpr uint32_t SPIN_LIMIT = 1u << 20;

__device__ inline void st_tag(unsigned long long* p, uint32_t tag, uint32_t v) {
    __hip_atomic_store(p, ((unsigned long long)tag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_tag(const unsigned long long* p) {
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint32_t hsh(uint32_t t, uint32_t s, uint32_t k) {
    uint32_t x = t * 2654435761u ^ (s * 40503u + k * 97u + 1u);
    x ^= x >> 13; x *= 0x5bd1e995u; x ^= x >> 15;
    return x & 1023u;
}
__device__ inline uint32_t flip_of(uint32_t t) { return ((t * 2654435761u) >> 7) & 1u; }

__device__ inline float work(float x, float y, int iters) {
#pragma unroll 8
    for (int i = 0; i < iters; ++i) x = __builtin_fmaf(x, 1.0000001f, y);
    return x;
}

template <int NW, bool LB, bool TICKET, int TPL>
__global__ __launch_bounds__(NW * 64) void k_probe(uint32_t n_tiles, unsigned long long* rec, uint32_t epoch,
                                                   int itA, int itC, uint32_t* ticket, uint32_t* abort_flag,
                                                   uint32_t* out, float* sink) {
    __shared__ float pad[NW * 1150];   // LDS like the real tile (~18.4 KB per 4 waves)
    __shared__ uint32_t s_pre[1 + NCNT];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x, ln = tid & 63u;
    uint32_t tile;
    if (TICKET) {
        if (tid == 0) s_tile = atomicAdd(ticket, 1u);
        __syncthreads();
        tile = s_tile;
    } else {
        tile = blockIdx.x;
    }
    float x = work((float)tid, (float)tile * 1e-9f, itA);
    pad[tid] = x;
    __syncthreads();
    if (LB) {
        if (tid < 64) {
            unsigned long long* me = rec + (size_t)tile * REC;
            // aggregate
            if (ln < NSTATE * NCNT) st_tag(me + ln, epoch, hsh(tile, ln / NCNT, ln % NCNT));
            if (ln == 8) st_tag(me + 8, epoch, flip_of(tile));
            uint32_t s_ex = 0, c_ex = 0;   // lane k of c_ex: counter k (lanes 0..3)
            if (tile > 0) {
                uint32_t polls = 0;
                while (true) {
                    unsigned long long wi[TPL], wa[TPL];
                    uint64_t incm[TPL], aggm[TPL], flips[TPL];
                    const unsigned long long* r[TPL];
#pragma unroll
                    for (int q = 0; q < TPL; ++q) {
                        const int64_t j = (int64_t)tile - 1 - (int64_t)ln - 64 * q;
                        const bool valid = j >= 0;
                        r[q] = rec + (size_t)(valid ? j : 0) * REC;
                        wi[q] = valid ? ld_tag(r[q] + 13) : 0ULL;
                        wa[q] = valid ? ld_tag(r[q] + 8) : 0ULL;
                    }
                    int qk = -1;
                    bool ready = true;
#pragma unroll
                    for (int q = 0; q < TPL; ++q) {
                        const int64_t j = (int64_t)tile - 1 - (int64_t)ln - 64 * q;
                        const bool valid = j >= 0;
                        incm[q] = __ballot(valid && (uint32_t)(wi[q] >> 32) == epoch);
                        aggm[q] = __ballot(valid && (uint32_t)(wa[q] >> 32) == epoch);
                        flips[q] = __ballot(valid && (uint32_t)(wa[q] >> 32) == epoch && ((uint32_t)wa[q] & 1u));
                    }
                    uint64_t need[TPL];
#pragma unroll
                    for (int q = 0; q < TPL; ++q) {
                        if (qk >= 0) { need[q] = 0; continue; }
                        if (incm[q]) {
                            qk = q;
                            const uint32_t k = __builtin_ctzll(incm[q]);
                            need[q] = k ? ((1ULL << k) - 1ULL) : 0ULL;
                        } else {
                            need[q] = ~0ULL;
                        }
                        if ((aggm[q] & need[q]) != need[q]) ready = false;
                    }
                    bool done = false;
                    if (qk >= 0 && ready) {
                        const uint32_t k = __builtin_ctzll(incm[qk]);
                        const uint32_t s_inc = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)wi[qk], (int)k);
                        // parity of flips strictly between lane's tile and the INC tile
                        uint32_t par_tot = 0;
#pragma unroll
                        for (int q = 0; q < TPL; ++q) par_tot += __popcll(flips[q] & need[q]);
                        uint32_t v[NCNT] = {0, 0, 0, 0};
                        bool ok = true;
#pragma unroll
                        for (int q = 0; q < TPL; ++q) {
                            if (q > qk) break;
                            // tiles after (nearer than) this lane's tile in round q, and in rounds < q
                            const uint64_t above = (ln + 1 < 64 ? (~0ULL << (ln + 1)) : 0ULL) & need[q];
                            uint32_t par = __popcll(flips[q] & above);
                            // flips in rounds q+1..qk (older tiles, between)
                            for (int p = q + 1; p < TPL; ++p) if (p <= qk) par += __popcll(flips[p] & need[p]);
                            const uint32_t ent = s_inc ^ (par & 1u);
                            const bool mine = (need[q] >> ln) & 1ULL;
                            const bool isinc = q == qk && ln == k;
#pragma unroll
                            for (int c = 0; c < NCNT; ++c) {
                                unsigned long long w = 0;
                                if (mine) w = ld_tag(r[q] + ent * NCNT + c);
                                else if (isinc) w = ld_tag(r[q] + 9 + c);
                                if ((mine || isinc) && (uint32_t)(w >> 32) != epoch) ok = false;
                                v[c] += (mine || isinc) ? (uint32_t)w : 0u;
                            }
                        }
                        if (__ballot(!ok) == 0) {
                            s_ex = s_inc ^ (par_tot & 1u);
                            for (int c = 0; c < NCNT; ++c) {
                                uint32_t sum = v[c];
                                for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
                                if ((int)ln == c) c_ex = sum;
                            }
                            done = true;
                        }
                    }
                    if (done) break;
                    if (++polls > SPIN_LIMIT || __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        if (ln == 0) __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        s_ex = 0xFFFF;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            // inclusive = exclusive then own aggregate for state s_ex
            const uint32_t se = s_ex & 1u;
            if (ln < NCNT) st_tag(me + 9 + ln, epoch, c_ex + hsh(tile, se, ln));
            if (ln == 0) st_tag(me + 13, epoch, se ^ flip_of(tile));
            if (ln == 0) s_pre[0] = s_ex;
            if (ln < NCNT) s_pre[1 + ln] = c_ex;
        }
        __syncthreads();
        if (tid <= NCNT) out[(size_t)tile * 8 + tid] = s_pre[tid];
    }
    float y = work(pad[(tid + 1) % (NW * 64)], x, itC);
    if (y == 12345.678f) sink[tile] = y;
}

template <int NW, bool LB, bool TICKET, int TPL>
float run(uint32_t n_tiles, int itA, int itC, unsigned long long* rec, uint32_t& epoch, uint32_t* ticket,
          uint32_t* abort_flag, uint32_t* out, float* sink, int reps, bool check) {
    hipEvent_t a, b;
    CK(hipEventCreate(&a)); CK(hipEventCreate(&b));
    float best = 1e9;
    for (int r = 0; r < reps; ++r) {
        ++epoch;
        CK(hipMemset(ticket, 0, 4));
        CK(hipEventRecord(a));
        hipLaunchKernelGGL((k_probe<NW, LB, TICKET, TPL>), dim3(n_tiles), dim3(NW * 64), 0, 0, n_tiles, rec, epoch, itA, itC,
                           ticket, abort_flag, out, sink);
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        float ms; CK(hipEventElapsedTime(&ms, a, b));
        if (ms < best) best = ms;
    }
    uint32_t ab = 0;
    CK(hipMemcpy(&ab, abort_flag, 4, hipMemcpyDeviceToHost));
    if (ab) { printf("  ABORT flag set (NW=%d LB=%d T=%d)\n", NW, LB, TICKET); CK(hipMemset(abort_flag, 0, 4)); }
    if (check && LB) {
        std::vector<uint32_t> h((size_t)n_tiles * 8);
        CK(hipMemcpy(h.data(), out, h.size() * 4, hipMemcpyDeviceToHost));
        // host recompute
        auto hs = [](uint32_t t, uint32_t s, uint32_t k) {
            uint32_t x = t * 2654435761u ^ (s * 40503u + k * 97u + 1u);
            x ^= x >> 13; x *= 0x5bd1e995u; x ^= x >> 15;
            return x & 1023u;
        };
        uint32_t s = 0, c[NCNT] = {0, 0, 0, 0};
        size_t bad = 0;
        for (uint32_t t = 0; t < n_tiles; ++t) {
            bool okk = h[(size_t)t * 8] == s;
            for (int q = 0; q < NCNT; ++q) okk = okk && h[(size_t)t * 8 + 1 + q] == c[q];
            if (!okk && bad++ < 3) printf("  mismatch tile %u: got s=%u c0=%u want s=%u c0=%u\n", t, h[(size_t)t*8], h[(size_t)t*8+1], s, c[0]);
            for (int q = 0; q < NCNT; ++q) c[q] += hs(t, s, q);
            s ^= ((t * 2654435761u) >> 7) & 1u;
        }
        printf("  check NW=%d ticket=%d: %zu bad tiles of %u\n", NW, TICKET, bad, n_tiles);
    }
    return best;
}

template <int NW>
void sweep(unsigned long long* rec, uint32_t& epoch, uint32_t* ticket, uint32_t* ab, uint32_t* out, float* sink,
           int perwaveA, int perwaveC) {
    const uint32_t n_tiles = (uint32_t)((1ULL << 30) / (NW * 2048));
    float base = run<NW, false, false, 1>(n_tiles, perwaveA, perwaveC, rec, epoch, ticket, ab, out, sink, 5, false);
    float l1 = run<NW, true, false, 1>(n_tiles, perwaveA, perwaveC, rec, epoch, ticket, ab, out, sink, 5, true);
    float l2 = run<NW, true, false, 2>(n_tiles, perwaveA, perwaveC, rec, epoch, ticket, ab, out, sink, 5, true);
    float l4 = run<NW, true, false, 4>(n_tiles, perwaveA, perwaveC, rec, epoch, ticket, ab, out, sink, 5, true);
    float lt = run<NW, true, true, 1>(n_tiles, perwaveA, perwaveC, rec, epoch, ticket, ab, out, sink, 5, true);
    printf("NW=%2d tile=%2d KiB tiles=%u: none %.3f ms | window 64 %.3f | 128 %.3f | 256 %.3f | 64 by ticket %.3f\n",
           NW, NW * 2, n_tiles, base, l1, l2, l4, lt);
    fflush(stdout);
}

int main(int argc, char** argv) {
    int A = argc > 1 ? atoi(argv[1]) : 300, C = argc > 2 ? atoi(argv[2]) : 650;
    const size_t max_tiles = (1ULL << 30) / (4 * 2048);
    unsigned long long* rec; uint32_t *ticket, *ab, *out; float* sink;
    CK(hipMalloc(&rec, max_tiles * REC * 8));
    CK(hipMemset(rec, 0, max_tiles * REC * 8));
    CK(hipMalloc(&ticket, 4)); CK(hipMalloc(&ab, 4)); CK(hipMemset(ab, 0, 4));
    CK(hipMalloc(&out, max_tiles * 32)); CK(hipMalloc(&sink, max_tiles * 4));
    uint32_t epoch = 0;
    printf("work per wave: A=%d C=%d VALU\n", A, C); fflush(stdout);
    sweep<4>(rec, epoch, ticket, ab, out, sink, A, C);
    sweep<8>(rec, epoch, ticket, ab, out, sink, A, C);
    sweep<16>(rec, epoch, ticket, ab, out, sink, A, C);
    return 0;
}
