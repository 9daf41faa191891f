// grid_barrier.hip -- what a grid-wide barrier costs against a kernel
// boundary on this GPU (a measurement behind the design of the FCG/V-cycle
// launches; not part of the library).  A kernel of `grid` workgroups (every
// `stride`-th one taking part) runs 200 barriers: a device-scope atomic
// arrival counter and a generation word polled with s_sleep, every wait
// bounded (a spin limit sets an error flag and the kernel exits).  Then 200
// back-to-back launches of a one-line kernel.  Build and run:
//   hipcc -O3 --offload-arch=gfx950 grid_barrier.hip -o grid_barrier && ./grid_barrier
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>

#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

struct Bar { unsigned count; unsigned gen; unsigned err; unsigned pad; };

// participants: blocks with (blockIdx.x % stride == 0); others exit at once
__device__ __forceinline__ void GridSync(Bar* b, unsigned n, unsigned& g) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned a = atomicAdd(&b->count, 1u);
    if (a == n - 1) {
      __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spins = 0;
      while (__hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++spins > (1u << 22)) { atomicOr(&b->err, 1u); break; }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
  }
  g += 1;
  __syncthreads();
}

__global__ void __launch_bounds__(256) KBarriers(Bar* b, int iters, int stride, float* data, int n) {
  if (blockIdx.x % stride) return;
  const unsigned parts = (gridDim.x + stride - 1) / stride;
  const unsigned me = blockIdx.x / stride;
  unsigned g = 0;
  for (int it = 0; it < iters; ++it) {
    // touch a little data: each participant writes 256 floats, reads a neighbour's
    const int k = (me * 256 + threadIdx.x) % n;
    const int k2 = (((me + 1) % parts) * 256 + threadIdx.x) % n;
    data[k] += 1.f;
    GridSync(b, parts, g);
    data[k] += data[k2] * 1e-9f;
    if (b->err) return;
  }
}

__global__ void __launch_bounds__(256) KTiny(float* data, int n) {
  const int k = (blockIdx.x * 256 + threadIdx.x) % n;
  data[k] += 1.f;
}

int main() {
  Bar* b; float* data; const int n = 1 << 20;
  CK(hipMalloc(&b, sizeof(Bar))); CK(hipMemset(b, 0, sizeof(Bar)));
  CK(hipMalloc(&data, n * 4)); CK(hipMemset(data, 0, n * 4));
  hipStream_t st; CK(hipStreamCreate(&st));
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  const int iters = 200;
  struct Cfg { int grid, stride; } cfgs[] = {{32, 1}, {64, 1}, {256, 1}, {256, 8}, {512, 8}, {1024, 1}};
  for (auto c : cfgs) {
    for (int rep = 0; rep < 2; ++rep) {
      CK(hipMemset(b, 0, sizeof(Bar)));
      CK(hipEventRecord(e0, st));
      KBarriers<<<c.grid, 256, 0, st>>>(b, iters, c.stride, data, n);
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1));
      Bar h; CK(hipMemcpy(&h, b, sizeof(Bar), hipMemcpyDeviceToHost));
      if (rep) printf("barrier grid %4d stride %d: %.2f us per barrier (err %u, gen %u)\n", c.grid, c.stride, ms * 1e3 / iters, h.err, h.gen);
    }
  }
  for (int grid : {32, 256, 2048}) {
    for (int rep = 0; rep < 2; ++rep) {
      CK(hipEventRecord(e0, st));
      for (int i = 0; i < iters; ++i) KTiny<<<grid, 256, 0, st>>>(data, n);
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1));
      if (rep) printf("tiny kernels grid %4d: %.2f us per launch\n", grid, ms * 1e3 / iters);
    }
  }
  printf("done\n");
  return 0;
}
