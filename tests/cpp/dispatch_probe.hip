// dispatch_probe -- what a dependent kernel boundary costs on this GPU, and
// what finishing a dependent phase inside the producing kernel costs instead
// (measurement infrastructure for the step kernels' design; run it under
// rocprofv3 --kernel-trace and read the per-kernel durations).
//
// Launched back to back on one stream, 1,000 times each:
//   k_heavy    4,096 one-wave blocks writing 8 MB and spinning (~140 us):
//              leaves the caches dirty, like k_cell_act
//   k_empty    161 blocks of 256 threads (the scan's grid at 40,962 cells)
//              that do nothing; k_empty1 the same with one wave
//   k_copy     one load and one store per cell (40,962 cells)
//   k_chain    two dependent loads per cell
//   k_prod     4,096 one-wave blocks: 2 KB of writes each, ten per-cell
//              outputs, a short spin
//   k_cons     161 blocks reading the per-cell outputs (a scan-sized pass)
//   k_fused_seq, k_fused_rel
//              k_prod's work, then in the same kernel: each block fences
//              (seq_cst resp. release, agent scope) and counts itself done;
//              the last 161 finishers wait for all (bounded spin), fence
//              (acquire) and do k_cons's work.  On a GPU whose L2 is split
//              over several dies the agent-scope fence writes back the die's
//              L2, once per block.
// Exit status 1 if a bounded wait ran out (a flag, never a hang).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kBlocks = 4096, kCells = 40962, kTiles = 161;

__global__ void k_heavy(uint32_t *buf, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = i; k < n; k += gridDim.x * blockDim.x) buf[k] = buf[k] * 3u + 1u;
    uint32_t x = (uint32_t)i;
    for (int j = 0; j < 2000; j++) x = x * 1664525u + 1013904223u;
    if (x == 0x12345678u) buf[0] = x;
}

__global__ void k_empty(uint32_t *buf) {
    if (threadIdx.x >= blockDim.x) buf[0] = 1u;   // never: keeps the argument live
}

__global__ void k_copy(const uint32_t *a, uint32_t *b, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i] + 1u;
}

__global__ void k_chain(const uint32_t *a, uint32_t *b, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[a[i] % (uint32_t)n] + 1u;
}

__device__ __forceinline__ void produce(uint32_t *big, uint32_t *cell, int b) {
    uint32_t *p = big + (size_t)b * 512;
    for (int k = threadIdx.x; k < 512; k += 64) p[k] = p[k] * 3u + (uint32_t)b;
    for (int k = threadIdx.x; k < 10; k += 64) cell[(b * 10 + k) % kCells] = (uint32_t)(b + k);
    uint32_t x = (uint32_t)(b * 64 + threadIdx.x);
    for (int j = 0; j < 3000; j++) x = x * 1664525u + 1013904223u;
    if (x == 0x12345678u) big[0] = x;
}

__device__ __forceinline__ void consume(const uint32_t *cell, uint32_t *out, int tile) {
    for (int k = threadIdx.x; k < 256; k += blockDim.x) {
        const int c = tile * 256 + k;
        if (c < kCells) out[c] = cell[c] + 1u;
    }
}

__global__ void k_prod(uint32_t *big, uint32_t *cell) { produce(big, cell, blockIdx.x); }
__global__ void k_cons(const uint32_t *cell, uint32_t *out) { consume(cell, out, blockIdx.x); }

template <bool kSeq>
__global__ void k_fused(uint32_t *big, uint32_t *cell, uint32_t *out, uint32_t *ctr, uint32_t *err) {
    produce(big, cell, blockIdx.x);
    __shared__ uint32_t t;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (kSeq) __threadfence();
        else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        t = atomicAdd(ctr, 1u);
    }
    __syncthreads();
    const uint32_t tk = t;
    if (tk < (uint32_t)(kBlocks - kTiles)) return;
    if (threadIdx.x == 0) {
        uint32_t spins = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < (uint32_t)kBlocks) {
            if (++spins > (1u << 22)) {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (kSeq) __threadfence();
        else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    consume(cell, out, (int)(tk - (uint32_t)(kBlocks - kTiles)));
}

__global__ void k_reset(uint32_t *ctr) { *ctr = 0u; }

#define CK(x)                                                               \
    do {                                                                    \
        if ((x) != hipSuccess) {                                            \
            std::fprintf(stderr, "dispatch_probe: %s failed\n", #x);        \
            return 2;                                                       \
        }                                                                   \
    } while (0)

}  // namespace

int main() {
    const int nheavy = 2 << 20;
    uint32_t *big, *cell, *out, *ctr, *err;
    CK(hipMalloc(&big, (size_t)kBlocks * 512 * 4));
    CK(hipMalloc(&cell, (size_t)kCells * 4));
    CK(hipMalloc(&out, (size_t)kCells * 4));
    CK(hipMalloc(&ctr, 256));
    CK(hipMalloc(&err, 4));
    CK(hipMemset(big, 0, (size_t)kBlocks * 512 * 4));
    CK(hipMemset(cell, 0, (size_t)kCells * 4));
    CK(hipMemset(ctr, 0, 256));
    CK(hipMemset(err, 0, 4));
    hipStream_t s;
    CK(hipStreamCreate(&s));
    for (int r = 0; r < 1000; r++) {
        hipLaunchKernelGGL(k_heavy, dim3(kBlocks), dim3(64), 0, s, big, nheavy);
        hipLaunchKernelGGL(k_empty, dim3(kTiles), dim3(256), 0, s, big);
        hipLaunchKernelGGL(k_copy, dim3(kTiles), dim3(256), 0, s, (const uint32_t *)cell, out, kCells);
        hipLaunchKernelGGL(k_chain, dim3(kTiles), dim3(256), 0, s, (const uint32_t *)cell, out, kCells);
        hipLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, s, big);
        hipLaunchKernelGGL(k_prod, dim3(kBlocks), dim3(64), 0, s, big, cell);
        hipLaunchKernelGGL(k_cons, dim3(kTiles), dim3(256), 0, s, (const uint32_t *)cell, out);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, ctr);
        hipLaunchKernelGGL(k_fused<true>, dim3(kBlocks), dim3(64), 0, s, big, cell, out, ctr, err);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, ctr);
        hipLaunchKernelGGL(k_fused<false>, dim3(kBlocks), dim3(64), 0, s, big, cell, out, ctr, err);
    }
    CK(hipStreamSynchronize(s));
    uint32_t e = 0;
    CK(hipMemcpy(&e, err, 4, hipMemcpyDeviceToHost));
    std::printf("dispatch_probe done, wait flag %u\n", e);
    return e ? 1 : 0;
}
