// learned_net.hip -- learned preconditioner: V-cycle + U-Net correction,
// solver-generated training pairs (see learned_net.h).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "cg.h"
#include "device_common.h"
#include "learned_net.h"

#define HG_CHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") +            \
                                                   hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace hg {

namespace {

__device__ __forceinline__ bool Done(const int* d) { return d && *d; }

inline unsigned Grid(long n) { return unsigned(std::min<long>((n + 255) / 256, 8192)); }

// s = D^-1 (r - A z) with the FP32 face coefficients of the finest level.
__global__ void __launch_bounds__(kTileThreads)
KPost(int nx, int ny, TileGrid tg, const float* __restrict__ kx, const float* __restrict__ ky,
      const float* __restrict__ dinv, const double* __restrict__ r, const float* __restrict__ z,
      float* __restrict__ s, const int* done) {
  if (Done(done)) return;
  int i, j;
  if (!TileCell(tg, nx, ny, i, j)) return;
  const int c = i + nx * j;
  const float kw = kx[i + (nx + 1) * j], ke = kx[i + 1 + (nx + 1) * j];
  const float ks = ky[i + nx * j], kn = ky[i + nx * (j + 1)];
  float v = (((kw + ke) + ks) + kn) * z[c];
  if (i > 0) v -= kw * z[c - 1];
  if (i < nx - 1) v -= ke * z[c + 1];
  if (j > 0) v -= ks * z[c - nx];
  if (j < ny - 1) v -= kn * z[c + nx];
  s[c] = dinv[c] * (float(r[c]) - v);
}

// r = b - A x (FP64 face coefficients).
__global__ void __launch_bounds__(256)
KResid64(int nx, int ny, const double* __restrict__ kx, const double* __restrict__ ky,
         const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r) {
  const long n = long(nx) * ny;
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    const int i = int(c % nx), j = int(c / nx);
    const double kw = kx[i + long(nx + 1) * j], ke = kx[i + 1 + long(nx + 1) * j];
    const double ks = ky[i + long(nx) * j], kn = ky[i + long(nx) * (j + 1)];
    double v = (((kw + ke) + ks) + kn) * x[c];
    if (i > 0) v -= kw * x[c - 1];
    if (i < nx - 1) v -= ke * x[c + 1];
    if (j > 0) v -= ks * x[c - nx];
    if (j < ny - 1) v -= kn * x[c + nx];
    r[c] = b[c] - v;
  }
}

// 64-bit mixing function (the MurmurHash3 finaliser).
__device__ __forceinline__ uint64_t Mix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// Multi-scale noise: sum over scales l = 0..6 of piecewise-constant values
// uniform in [-g^l, g^l) on 2^l x 2^l blocks (g = 1: equal amplitudes;
// g = 2: equal energy-norm contributions for a Laplacian-like operator).
__global__ void __launch_bounds__(256)
KNoise(int nx, int ny, uint64_t seed, double g, double* __restrict__ b) {
  const long n = long(nx) * ny;
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    const int i = int(c % nx), j = int(c / nx);
    double v = 0., amp = 1.;
    for (int l = 0; l <= 6; ++l, amp *= g) {
      const uint64_t key = seed * 0x9e3779b97f4a7c15ULL + (uint64_t(l) << 58) +
                           (uint64_t(unsigned(j >> l)) << 29) + uint64_t(unsigned(i >> l));
      v += amp * (double(Mix(key) >> 11) * (2.0 / 9007199254740992.0) - 1.0);
    }
    b[c] = v;
  }
}

// s = D^-1 r (FP32)
__global__ void KScaleDinv(long n, const float* __restrict__ dinv, const double* __restrict__ r,
                           float* __restrict__ s, const int* done) {
  if (Done(done)) return;
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L)
    s[c] = dinv[c] * float(r[c]);
}

// s = D^-1/2 r (FP32)
__global__ void KScaleSqrtDinv(long n, const float* __restrict__ dinv,
                               const double* __restrict__ r, float* __restrict__ s,
                               const int* done) {
  if (Done(done)) return;
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L)
    s[c] = sqrtf(dinv[c]) * float(r[c]);
}

// z = keep * z + D^-1/2 y
__global__ void KAddScaled(long n, const float* __restrict__ dinv, const float* __restrict__ y,
                           int keep, float* __restrict__ z, const int* done) {
  if (Done(done)) return;
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    const float v = sqrtf(dinv[c]) * y[c];
    z[c] = keep ? z[c] + v : v;
  }
}

__global__ void KSqrt(long n, const float* __restrict__ a, float* __restrict__ out) {
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L)
    out[c] = sqrtf(a[c]);
}

// t = e (FP32)
__global__ void KToF32(long n, const double* __restrict__ e, float* __restrict__ t) {
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) t[c] = float(e[c]);
}

// Per-block (sum, sum of squares).
template <class T>
__global__ void __launch_bounds__(256)
KSums(long n, const T* __restrict__ a, double* __restrict__ part) {
  double s = 0., q = 0.;
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    const double v = double(a[c]);
    s += v;
    q += v * v;
  }
  s = WaveSum(s);
  q = WaveSum(q);
  __shared__ double red[2][4];
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s;
    red[1][threadIdx.x >> 6] = q;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// Per-block t.(A t) with the FP32 face coefficients.
__global__ void __launch_bounds__(256)
KEnergyDot(int nx, int ny, const float* __restrict__ kx, const float* __restrict__ ky,
           const float* __restrict__ t, double* __restrict__ part) {
  const long n = long(nx) * ny;
  double s = 0.;
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    const int i = int(c % nx), j = int(c / nx);
    const float kw = kx[i + long(nx + 1) * j], ke = kx[i + 1 + long(nx + 1) * j];
    const float ks = ky[i + long(nx) * j], kn = ky[i + long(nx) * (j + 1)];
    double q = double(((kw + ke) + ks) + kn) * t[c];
    if (i > 0) q -= double(kw) * t[c - 1];
    if (i < nx - 1) q -= double(ke) * t[c + 1];
    if (j > 0) q -= double(ks) * t[c - nx];
    if (j < ny - 1) q -= double(kn) * t[c + nx];
    s += double(t[c]) * q;
  }
  s = WaveSum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[2 * blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void KShift64(long n, double* __restrict__ a, double shift) {
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) a[c] -= shift;
}

// t = (xs - x) - z
__global__ void KTarget(long n, const double* __restrict__ xs, const double* __restrict__ x,
                        const float* __restrict__ z, float* __restrict__ t) {
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L)
    t[c] = float((xs[c] - x[c]) - double(z[c]));
}

// t = (t - mean) * scale; s *= scale
__global__ void KNormalise(long n, float* __restrict__ t, float* __restrict__ s, float mean,
                           float scale) {
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    t[c] = (t[c] - mean) * scale;
    s[c] *= scale;
  }
}

__global__ void KAddZ(long n, double* __restrict__ x, const float* __restrict__ z) {
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) x[c] += double(z[c]);
}

constexpr int kSumBlocks = 512;

// t.(A t) for the FP32 face-form operator.
double EnergyDot(int nx, int ny, const float* kx, const float* ky, const float* t, double* part,
                 double* host, hipStream_t st) {
  KEnergyDot<<<kSumBlocks, 256, 0, st>>>(nx, ny, kx, ky, t, part);
  HG_CHECK(hipMemcpyAsync(host, part, 2 * kSumBlocks * 8, hipMemcpyDeviceToHost, st));
  HG_CHECK(hipStreamSynchronize(st));
  double s = 0.;
  for (int b = 0; b < kSumBlocks; ++b) s += host[2 * b];
  return s;
}

// Mean and variance of a[0..n).
template <class T>
void Moments(long n, const T* a, double* part, double* host, hipStream_t st, double* mean,
             double* var) {
  KSums<T><<<kSumBlocks, 256, 0, st>>>(n, a, part);
  HG_CHECK(hipMemcpyAsync(host, part, 2 * kSumBlocks * 8, hipMemcpyDeviceToHost, st));
  HG_CHECK(hipStreamSynchronize(st));
  double s = 0., q = 0.;
  for (int b = 0; b < kSumBlocks; ++b) {
    s += host[2 * b];
    q += host[2 * b + 1];
  }
  *mean = s / double(n);
  *var = std::max(0., q / double(n) - (*mean) * (*mean));
}

struct Dev {
  void* p = nullptr;
  explicit Dev(size_t bytes) { HG_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace

TrainOpts TrainOptsFromAbi(const hg_train_opts& a, bool project) {
  if (a.rhs_count < 1 || a.mg_iters < 1 || a.steps < 0 || !(a.lr > 0.f) ||
      long(a.rhs_count) * a.mg_iters > 4096)
    throw std::invalid_argument(
        "train options: rhs_count >= 1, mg_iters >= 1, rhs_count*mg_iters <= 4096, steps >= 0, "
        "lr > 0");
  TrainOpts o;
  o.rhs_count = a.rhs_count;
  o.mg_iters = a.mg_iters;
  o.steps = a.steps;
  o.lr = a.lr;
  o.seed = a.seed;
  o.project = project;
  o.linear = (a.flags & HG_TRAIN_LINEAR) != 0;
  o.tied = (a.flags & HG_TRAIN_TIED) != 0;
  o.standalone = (a.flags & HG_TRAIN_STANDALONE) != 0;
  return o;
}

LearnedNet::LearnedNet(int nx, int ny, MgPrecond* mg) : nx_(nx), ny_(ny), mg_(mg) {
  if (!Supported(nx, ny)) throw std::runtime_error("LearnedNet: unsupported grid");
  net_.reset(new UNet(nx, ny));
  net_->SetMetric(mg->level(0).kx, mg->level(0).ky);
  HG_CHECK(hipMalloc(&s_, size_t(nx) * ny * 4));
  HG_CHECK(hipMalloc(&y_, size_t(nx) * ny * 4));
  HG_CHECK(hipMalloc(&dsq_, size_t(nx) * ny * 4));
  net_->InitWeights(1, 0);
}

LearnedNet::~LearnedNet() {
  for (float* p : {s_, y_, dsq_})
    if (p) (void)hipFree(p);
}

void LearnedNet::Apply(const double* r, float* z, hipStream_t st, const int* d_done) {
  const MgLevel& L = mg_->level(0);
  const long n = long(nx_) * ny_;
  if (net_->tied()) {
    // M r = [V r +] S N(S r), S = D^-1/2
    KScaleSqrtDinv<<<Grid(n), 256, 0, st>>>(n, L.dinv, r, s_, d_done);
    net_->Forward(s_, y_, nullptr, st, d_done);
    if (!standalone_) mg_->Apply(r, z, st, d_done);
    KAddScaled<<<Grid(n), 256, 0, st>>>(n, L.dinv, y_, standalone_ ? 0 : 1, z, d_done);
    return;
  }
  if (standalone_) {
    KScaleDinv<<<Grid(n), 256, 0, st>>>(n, L.dinv, r, s_, d_done);
    net_->Forward(s_, z, nullptr, st, d_done);
    return;
  }
  mg_->Apply(r, z, st, d_done);
  KPost<<<L.tg.blocks(), kTileThreads, 0, st>>>(nx_, ny_, L.tg, L.kx, L.ky, L.dinv, r, z, s_,
                                                 d_done);
  net_->Forward(s_, z, z, st, d_done);
}

TrainReport LearnedNet::Train(CgSolver& cg, const TrainOpts& o) {
  hipStream_t st = cg.stream();
  const auto t0 = std::chrono::steady_clock::now();
  const long n = long(nx_) * ny_;
  const int npairs = std::max(1, o.rhs_count * o.mg_iters);
  TrainReport rep;
  Dev b(n * 8), xs(n * 8), x(n * 8), r(n * 8), z(n * 4), part(2 * kSumBlocks * 8);
  Dev pairs(size_t(npairs) * 2 * n * 4);
  std::vector<double> host(2 * kSumBlocks);
  float* ps = pairs.as<float>();
  const MgLevel& L = mg_->level(0);
  net_->SetTied(o.tied);
  net_->SetLinear(o.linear || o.tied);
  standalone_ = o.standalone;
  KSqrt<<<Grid(n), 256, 0, st>>>(n, L.dinv, dsq_);
  net_->SetOutputScale(o.tied ? dsq_ : nullptr);
  net_->SetScaleFree(true);
  double mean = 0., var = 0.;
  // Each pair: input s_k, target t_k, both scaled so that t^T A t = n (the
  // loss is the relative energy-norm error) after the target is projected
  // to zero mean (singular operator).
  auto finish_pair = [&](float* s_k, float* t_k) {
    if (o.project) {
      Moments(n, t_k, part.as<double>(), host.data(), st, &mean, &var);
      KNormalise<<<Grid(n), 256, 0, st>>>(n, t_k, s_k, float(mean), 1.f);
    }
    const double tat = EnergyDot(nx_, ny_, L.kx, L.ky, t_k, part.as<double>(), host.data(), st);
    const float scale = tat > 0. ? float(std::sqrt(double(n) / tat)) : 0.f;
    KNormalise<<<Grid(n), 256, 0, st>>>(n, t_k, s_k, 0.f, scale);
  };
  int k = 0;
  if (o.tied || o.standalone) {
    // residuals r: multi-scale noise; e = A^-1 r from a tight FCG + V-cycle
    // solve; input S r (tied) or D^-1 r; target e (standalone) or e - V r
    for (; k < npairs; ++k) {
      float* s_k = ps + size_t(2 * k) * n;
      float* t_k = s_k + n;
      KNoise<<<Grid(n), 256, 0, st>>>(nx_, ny_, o.seed * 1000003ULL + uint64_t(k), 1.0,
                                      r.as<double>());
      if (o.project) {
        Moments(n, r.as<double>(), part.as<double>(), host.data(), st, &mean, &var);
        KShift64<<<Grid(n), 256, 0, st>>>(n, r.as<double>(), mean);
      }
      CgResult ref = cg.Solve(r.as<double>(), xs.as<double>(), 1e-10, 4000, 1, o.project, mg_);
      if (!ref.converged) throw std::runtime_error("LearnedNet::Train: reference solve failed");
      if (o.tied)
        KScaleSqrtDinv<<<Grid(n), 256, 0, st>>>(n, L.dinv, r.as<double>(), s_k, nullptr);
      else
        KScaleDinv<<<Grid(n), 256, 0, st>>>(n, L.dinv, r.as<double>(), s_k, nullptr);
      if (o.standalone) {
        KToF32<<<Grid(n), 256, 0, st>>>(n, xs.as<double>(), t_k);
      } else {
        mg_->Apply(r.as<double>(), z.as<float>(), st, nullptr);
        HG_CHECK(hipMemsetAsync(x.p, 0, n * 8, st));
        KTarget<<<Grid(n), 256, 0, st>>>(n, xs.as<double>(), x.as<double>(), z.as<float>(), t_k);
      }
      finish_pair(s_k, t_k);
    }
  } else {
    // stationary V-cycle iteration x_{k+1} = x_k + V(b - A x_k) on random
    // right-hand sides: input D^-1 (r_k - A V r_k), target e_k - V r_k
    for (int q = 0; q < o.rhs_count; ++q) {
      KNoise<<<Grid(n), 256, 0, st>>>(nx_, ny_, o.seed * 1000003ULL + uint64_t(q), 1.0,
                                      b.as<double>());
      if (o.project) {
        Moments(n, b.as<double>(), part.as<double>(), host.data(), st, &mean, &var);
        KShift64<<<Grid(n), 256, 0, st>>>(n, b.as<double>(), mean);
      }
      CgResult ref = cg.Solve(b.as<double>(), xs.as<double>(), 1e-11, 4000, 1, o.project, mg_);
      if (!ref.converged) throw std::runtime_error("LearnedNet::Train: reference solve failed");
      HG_CHECK(hipMemsetAsync(x.p, 0, n * 8, st));
      for (int it = 0; it < o.mg_iters; ++it, ++k) {
        float* s_k = ps + size_t(2 * k) * n;
        float* t_k = s_k + n;
        KResid64<<<Grid(n), 256, 0, st>>>(nx_, ny_, cg.kx(), cg.ky(), b.as<double>(),
                                          x.as<double>(), r.as<double>());
        mg_->Apply(r.as<double>(), z.as<float>(), st, nullptr);
        KPost<<<L.tg.blocks(), kTileThreads, 0, st>>>(nx_, ny_, L.tg, L.kx, L.ky, L.dinv,
                                                       r.as<double>(), z.as<float>(), s_k,
                                                       nullptr);
        KTarget<<<Grid(n), 256, 0, st>>>(n, xs.as<double>(), x.as<double>(), z.as<float>(), t_k);
        KAddZ<<<Grid(n), 256, 0, st>>>(n, x.as<double>(), z.as<float>());
        finish_pair(s_k, t_k);
      }
    }
  }
  rep.pairs = k;
  std::mt19937_64 rng(o.seed ^ 0x5eedULL);
  std::uniform_int_distribution<int> pick(0, k - 1);
  const int steps = std::max(0, o.steps);
  const int win = std::max(1, steps / 10);
  double first = 0., last = 0., alpha = 0.;
  const float lr_min = o.lr * 0.1f;
  for (int s = 0; s < steps; ++s) {
    const int idx = pick(rng);
    const float* s_k = ps + size_t(2 * idx) * n;
    const float lr = lr_min + 0.5f * (o.lr - lr_min) *
                                  (1.f + std::cos(3.14159265f * float(s) / float(steps)));
    double loss = 0.;
    net_->TrainStep(s_k, s_k + n, lr, st, &loss);
    if (!(loss == loss)) throw std::runtime_error("LearnedNet::Train: loss is NaN");
    if (s < win) first += loss;
    if (s >= steps - win) {
      last += loss;
      alpha += net_->last_alpha();
    }
  }
  rep.steps = steps;
  rep.loss_first = steps ? first / win : 0.;
  rep.loss_last = steps ? last / win : 0.;
  // the loss is scale-free: the output is multiplied by the mean best
  // multiple of the last steps
  alpha = steps ? alpha / win : 0.;
  rep.alpha = alpha;
  if (steps > 0 && alpha > 0.) net_->ScaleOutput(float(alpha), st);
  HG_CHECK(hipStreamSynchronize(st));
  // Acceptance on a held-out right-hand side, by the true FP64 residual:
  // FCG with this preconditioner against the method without it (FCG with
  // the V-cycle alone; Jacobi-PCG for a standalone net).
  trained_ = false;
  if (steps > 0 && alpha > 0.) {
    KNoise<<<Grid(n), 256, 0, st>>>(nx_, ny_, o.seed * 1000003ULL + 999983ULL, 1.0,
                                    b.as<double>());
    if (o.project) {
      Moments(n, b.as<double>(), part.as<double>(), host.data(), st, &mean, &var);
      KShift64<<<Grid(n), 256, 0, st>>>(n, b.as<double>(), mean);
    }
    const double tol = 1e-8;
    CgResult with = cg.Solve(b.as<double>(), x.as<double>(), tol, 20000, 1, o.project, this);
    CgResult without = o.standalone
                           ? cg.Solve(b.as<double>(), x.as<double>(), tol, 20000, 0, o.project)
                           : cg.Solve(b.as<double>(), x.as<double>(), tol, 20000, 1, o.project,
                                      mg_);
    rep.val_iters = with.iterations;
    rep.val_iters_baseline = without.iterations;
    rep.val_ms = with.ms;
    rep.val_ms_baseline = without.ms;
    // used from now on only when it is both fewer iterations and faster
    trained_ = with.converged && !with.used_fallback && with.iterations < without.iterations &&
               with.ms < without.ms;
  }
  rep.accepted = trained_;
  rep.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

}  // namespace hg
