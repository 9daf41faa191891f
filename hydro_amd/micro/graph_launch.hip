// graph_launch.hip -- per-kernel cost of 200 tiny kernels launched one by
// one against one hipGraph replay of the same 200, from 1 and 8 host
// threads, each on its own stream (a measurement behind the C5 batch
// design; not part of the library).  Build and run:
//   hipcc -O2 --offload-arch=gfx950 graph_launch.hip -o graph_launch -pthread && ./graph_launch
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdio>
#include <thread>
#include <utility>
#include <vector>

#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP %s at %d\n", hipGetErrorString(e), __LINE__); exit(1); } } while (0)

__global__ void KTiny(float* d, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) d[k] += 1.f;
}

constexpr int kK = 200, kReps = 20;

std::pair<double, double> RunThread(bool graph, int grid) {
  hipStream_t st; CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  float* d; const int n = grid * 256; CK(hipMalloc(&d, n * 4)); CK(hipMemsetAsync(d, 0, n * 4, st));
  hipGraphExec_t ex = nullptr;
  if (graph) {
    hipGraph_t g;
    CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < kK; ++i) KTiny<<<grid, 256, 0, st>>>(d, n);
    CK(hipStreamEndCapture(st, &g));
    CK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    CK(hipGraphDestroy(g));
  }
  CK(hipStreamSynchronize(st));
  auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < kReps; ++r) {
    if (graph) CK(hipGraphLaunch(ex, st));
    else for (int i = 0; i < kK; ++i) KTiny<<<grid, 256, 0, st>>>(d, n);
  }
  auto t1 = std::chrono::steady_clock::now();  // host enqueue done
  CK(hipStreamSynchronize(st));
  auto t2 = std::chrono::steady_clock::now();
  if (ex) CK(hipGraphExecDestroy(ex));
  CK(hipFree(d)); CK(hipStreamDestroy(st));
  const double enq = std::chrono::duration<double, std::micro>(t1 - t0).count() / (kReps * kK);
  const double tot = std::chrono::duration<double, std::micro>(t2 - t0).count() / (kReps * kK);
  return {enq, tot};
}

int main() {
  for (int grid : {8, 64}) {
    for (bool graph : {false, true}) {
      for (int nt : {1, 8}) {
        std::vector<std::pair<double, double>> res(nt);
        std::vector<std::thread> th;
        auto t0 = std::chrono::steady_clock::now();
        for (int t = 0; t < nt; ++t) th.emplace_back([&, t]() { res[t] = RunThread(graph, grid); });
        for (auto& x : th) x.join();
        auto t1 = std::chrono::steady_clock::now();
        const double wall = std::chrono::duration<double, std::micro>(t1 - t0).count();
        double enq = 0, tot = 0;
        for (auto& v : res) { enq += v.first; tot += v.second; }
        printf("grid %3d %-6s threads %d: host enqueue %.2f us/kernel, per-thread total %.2f us/kernel, "
               "wall %.2f us per kernel of all threads\n", grid, graph ? "graph" : "launch", nt,
               enq / nt, tot / nt, wall / (nt * kReps * kK));
      }
    }
  }
  printf("done\n");
  return 0;
}
