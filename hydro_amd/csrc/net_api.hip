// net_api.hip -- C-ABI of the learned preconditioner's network alone
// (hg_net_*): checks against the CPU restatement and throughput figures.

#include <algorithm>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hydro_gpu.h"
#include "net.h"

#define HG_CHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") +            \
                                                   hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace {

__global__ void KFillNoise(long n, unsigned seed, float* __restrict__ a) {
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    unsigned h = unsigned(c) * 2654435761u + seed;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    a[c] = float(h >> 8) * (2.f / 16777216.f) - 1.f;
  }
}

}  // namespace

struct hg_net {
  int nx = 0, ny = 0;
  hipStream_t st = nullptr;
  std::unique_ptr<hg::UNet> net;
  float *x = nullptr, *y = nullptr, *t = nullptr;
  ~hg_net() {
    if (st) (void)hipStreamSynchronize(st);
    net.reset();
    for (float* p : {x, y, t})
      if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
  size_t bytes() const { return size_t(nx) * ny * 4; }
};

namespace {

template <class F>
hg_status Run(const char* name, F&& fn) {
  try {
    fn();
    return HG_OK;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "%s: %s\n", name, ex.what());
    return HG_ERR_HIP;
  }
}

}  // namespace

extern "C" {

hg_status hg_net_create(int32_t nx, int32_t ny, hg_net** out) {
  if (!out) return HG_ERR_BAD_ARG;
  *out = nullptr;
  if (!hg::UNet::Supported(nx, ny)) return HG_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HG_ERR_HIP;
  return Run("hg_net_create", [&]() {
    std::unique_ptr<hg_net> h(new hg_net());
    h->nx = nx;
    h->ny = ny;
    HG_CHECK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    h->net.reset(new hg::UNet(nx, ny));
    HG_CHECK(hipMalloc(&h->x, h->bytes()));
    HG_CHECK(hipMalloc(&h->y, h->bytes()));
    HG_CHECK(hipMalloc(&h->t, h->bytes()));
    h->net->InitWeights(1, h->st);
    *out = h.release();
  });
}

void hg_net_destroy(hg_net* h) { delete h; }

int64_t hg_net_num_weights(const hg_net* h) { return h ? h->net->num_weights() : -1; }

hg_status hg_net_configure(hg_net* h, int32_t flags) {
  if (!h || (flags & ~(HG_TRAIN_LINEAR | HG_TRAIN_TIED))) return HG_ERR_BAD_ARG;
  h->net->SetTied((flags & HG_TRAIN_TIED) != 0);
  h->net->SetLinear((flags & (HG_TRAIN_LINEAR | HG_TRAIN_TIED)) != 0);
  return Run("hg_net_configure", [&]() {
    std::vector<float> w(h->net->num_weights());
    h->net->GetWeights(w.data(), h->st);
    h->net->SetWeights(w.data(), h->st);  // refreshes the derived (tied) weights
  });
}

hg_status hg_net_set_weights(hg_net* h, const float* w) {
  if (!h || !w) return HG_ERR_BAD_ARG;
  return Run("hg_net_set_weights", [&]() { h->net->SetWeights(w, h->st); });
}

hg_status hg_net_get_weights(hg_net* h, float* w) {
  if (!h || !w) return HG_ERR_BAD_ARG;
  return Run("hg_net_get_weights", [&]() { h->net->GetWeights(w, h->st); });
}

hg_status hg_net_forward(hg_net* h, const float* x, float* y) {
  if (!h || !x || !y) return HG_ERR_BAD_ARG;
  return Run("hg_net_forward", [&]() {
    HG_CHECK(hipMemcpyAsync(h->x, x, h->bytes(), hipMemcpyHostToDevice, h->st));
    h->net->Forward(h->x, h->y, nullptr, h->st, nullptr);
    HG_CHECK(hipMemcpyAsync(y, h->y, h->bytes(), hipMemcpyDeviceToHost, h->st));
    HG_CHECK(hipStreamSynchronize(h->st));
  });
}

hg_status hg_net_gradient(hg_net* h, const float* x, const float* gy, float* gw) {
  if (!h || !x || !gy || !gw) return HG_ERR_BAD_ARG;
  return Run("hg_net_gradient", [&]() {
    HG_CHECK(hipMemcpyAsync(h->x, x, h->bytes(), hipMemcpyHostToDevice, h->st));
    HG_CHECK(hipMemcpyAsync(h->t, gy, h->bytes(), hipMemcpyHostToDevice, h->st));
    h->net->Forward(h->x, h->y, nullptr, h->st, nullptr);
    h->net->ZeroGrad(h->st);
    h->net->Backward(h->x, h->t, h->st);
    h->net->GetGrad(gw, h->st);
  });
}

hg_status hg_net_train_step(hg_net* h, const float* x, const float* t, float lr, double* loss) {
  if (!h || !x || !t || !(lr > 0.f)) return HG_ERR_BAD_ARG;
  return Run("hg_net_train_step", [&]() {
    HG_CHECK(hipMemcpyAsync(h->x, x, h->bytes(), hipMemcpyHostToDevice, h->st));
    HG_CHECK(hipMemcpyAsync(h->t, t, h->bytes(), hipMemcpyHostToDevice, h->st));
    double l = 0.;
    h->net->TrainStep(h->x, h->t, lr, h->st, &l);
    if (loss) *loss = l;
  });
}

hg_status hg_net_time(hg_net* h, int32_t reps, double* forward_ms, double* train_ms,
                      double* forward_flops, double* train_flops) {
  if (!h || reps < 1) return HG_ERR_BAD_ARG;
  return Run("hg_net_time", [&]() {
    const long n = long(h->nx) * h->ny;
    const unsigned g = unsigned(std::min<long>((n + 255) / 256, 4096));
    KFillNoise<<<g, 256, 0, h->st>>>(n, 7u, h->x);
    KFillNoise<<<g, 256, 0, h->st>>>(n, 11u, h->t);
    // the weights are restored afterwards: the timed training steps change them
    std::vector<float> saved(h->net->num_weights());
    h->net->GetWeights(saved.data(), h->st);
    hipEvent_t e[3];
    for (auto& ev : e) HG_CHECK(hipEventCreate(&ev));
    h->net->Forward(h->x, h->y, nullptr, h->st, nullptr);  // warm-up
    h->net->TrainStep(h->x, h->t, 1e-4f, h->st, nullptr);
    HG_CHECK(hipEventRecord(e[0], h->st));
    for (int r = 0; r < reps; ++r) h->net->Forward(h->x, h->y, nullptr, h->st, nullptr);
    HG_CHECK(hipEventRecord(e[1], h->st));
    for (int r = 0; r < reps; ++r) h->net->TrainStep(h->x, h->t, 1e-4f, h->st, nullptr);
    HG_CHECK(hipEventRecord(e[2], h->st));
    HG_CHECK(hipEventSynchronize(e[2]));
    float a = 0.f, b = 0.f;
    HG_CHECK(hipEventElapsedTime(&a, e[0], e[1]));
    HG_CHECK(hipEventElapsedTime(&b, e[1], e[2]));
    for (auto& ev : e) (void)hipEventDestroy(ev);
    h->net->SetWeights(saved.data(), h->st);
    if (forward_ms) *forward_ms = a / reps;
    if (train_ms) *train_ms = b / reps;
    if (forward_flops) *forward_flops = h->net->ForwardFlops();
    if (train_flops) *train_flops = h->net->TrainFlops();
  });
}

}  // extern "C"
