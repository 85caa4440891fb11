// vk_fft.hip -- the batched 2D real FFT over the periodic x-z planes of
// P2S/S2P (p2s_fftw2d.f), written for gfx950.
//
// Forward (P2S): per row (j,k) a real FFT of length nx along x, done as a
// complex FFT of length nx/2 on the packed even/odd samples plus the
// split post-processing, written straight into the spectrum layout
// PHIT(0:nx/2, j, k); then per column (I, j) a complex FFT of length nz
// along z in place, with the reference's division by nx*nz.  Inverse
// (S2P): the z inverse in place, then the complex-to-real x pass with the
// reference's wrap copies.  Rows and columns are staged in LDS; the FFTs
// are Stockham autosort passes of radix 4, 3 and 2 (sizes 2^a 3^b), with
// twiddles read from a table computed on the host in extended precision.
#include <cmath>
#include <map>

#include "vk_kernels.h"
#include "vk_runtime.h"

namespace vk {

struct FftPlan {
  int n = 0;
  int nst = 0;
  int r[20] = {0};
  const double2* tw = nullptr;  // e^{-2 pi i m / n}, m = 0..n-1
};

namespace {
std::map<int, FftPlan>& plans() {
  static std::map<int, FftPlan> p;
  return p;
}
std::map<int, const double2*>& tables() {
  static std::map<int, const double2*> t;
  return t;
}
}  // namespace

// e^{-2 pi i m / n} for m = 0..n-1, built once per n.
const double2* twiddle_table(int n) {
  auto it = tables().find(n);
  if (it != tables().end()) return it->second;
  std::vector<double2> h(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; m++) {
    long double ang = -2.0L * pi * (long double)m / (long double)n;
    h[m] = make_double2((double)cosl(ang), (double)sinl(ang));
  }
  double2* d = nullptr;
  if (hipMalloc(&d, sizeof(double2) * n) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  tables()[n] = d;
  return d;
}

// Factor n into radix 4, 3, 2 passes; false if another prime divides n.
static bool fft_plan(int n, FftPlan& p) {
  auto it = plans().find(n);
  if (it != plans().end()) {
    p = it->second;
    return true;
  }
  FftPlan q;
  q.n = n;
  int m = n;
  while (m % 4 == 0 && m > 1) {
    q.r[q.nst++] = 4;
    m /= 4;
  }
  while (m % 3 == 0) {
    q.r[q.nst++] = 3;
    m /= 3;
  }
  while (m % 2 == 0) {
    q.r[q.nst++] = 2;
    m /= 2;
  }
  if (m != 1) return false;
  q.tw = twiddle_table(n);
  if (!q.tw) return false;
  plans()[n] = q;
  p = q;
  return true;
}

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 conjd(double2 a) { return make_double2(a.x, -a.y); }
// multiply by sign * i
__device__ inline double2 mul_si(double2 a, int sign) {
  return sign < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// Stockham autosort FFT of length p.n in LDS, threads t = 0..T-1 of one
// group.  sign = -1 forward, +1 backward, unnormalised.  Returns the
// buffer holding the result (a or b).  All threads of the block must call
// it together (it contains block barriers).
__device__ double2* fft_lds(double2* a, double2* b, int t, int T, const FftPlan& p, int sign) {
  const int n = p.n;
  int ns = 1;
  double2* src = a;
  double2* dst = b;
  for (int s = 0; s < p.nst; s++) {
    const int R = p.r[s];
    const int nb = n / R;
    const int twstep = n / (ns * R);
    for (int j = t; j < nb; j += T) {
      const int kk = j % ns;
      double2 v[4];
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (r < R) {
          v[r] = src[j + r * nb];
          if (r > 0) {
            double2 w = p.tw[(r * kk * twstep) % n];
            if (sign > 0) w = conjd(w);
            v[r] = cmul(v[r], w);
          }
        }
      double2 o[4];
      if (R == 4) {
        double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        double2 s13 = cadd(v[1], v[3]), d13 = mul_si(csub(v[1], v[3]), sign);
        o[0] = cadd(s02, s13);
        o[2] = csub(s02, s13);
        o[1] = cadd(d02, d13);
        o[3] = csub(d02, d13);
      } else if (R == 3) {
        const double h = 0.86602540378443864676;  // sqrt(3)/2
        double2 sm = cadd(v[1], v[2]), df = csub(v[1], v[2]);
        o[0] = cadd(v[0], sm);
        double2 tt = make_double2(v[0].x - 0.5 * sm.x, v[0].y - 0.5 * sm.y);
        double2 rot = mul_si(make_double2(h * df.x, h * df.y), sign);
        o[1] = cadd(tt, rot);
        o[2] = csub(tt, rot);
      } else {
        o[0] = cadd(v[0], v[1]);
        o[1] = csub(v[0], v[1]);
      }
      const int base = (j / ns) * ns * R + kk;
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (r < R) dst[base + r * ns] = o[r];
    }
    __syncthreads();
    double2* tmp = src;
    src = dst;
    dst = tmp;
    ns *= R;
  }
  return src;
}

// Which sample feeds the forward x pass: a stored field, or the Poisson
// right-hand side div(u_temp)/dt computed on the fly (mkp.f:88-121, where
// index 0 of x and z is the periodic image of nx and nz).
struct SampleField {
  const double* fp;
  __device__ double operator()(const Geom& g, int i, int j, int k) const {
    return fp[g.at(i, j, k)];
  }
};
struct SampleDiv {
  const double *u, *v, *w, *dx, *dy, *dz;
  double dt;
  __device__ double operator()(const Geom& g, int i, int j, int k) const {
    int ii = (i == 0) ? g.nx : i, kk = (k == 0) ? g.nz : k;
    long long o = g.at(ii, j, kk);
    return ((u[o] - u[o - 1]) / dx[ii] + (v[o] - v[o - g.sy]) / dy[j] +
            (w[o] - w[o - g.sz]) / dz[kk]) /
           dt;
  }
};

constexpr int kRowsPerBlock = 4;
constexpr int kRowThreads = 64;

// forward x pass: rows (j, k) = (1 + blockIdx.x*4 + row, blockIdx.y)
template <typename Sample>
__global__ __launch_bounds__(kRowsPerBlock* kRowThreads) void k_rfft_x(Geom g, Sample smp,
                                                                        double* phit, FftPlan p,
                                                                        const double2* twl) {
  extern __shared__ double2 sh[];
  const int M = p.n;  // nx/2
  const int row = threadIdx.x / kRowThreads, t = threadIdx.x % kRowThreads;
  const int j = 1 + blockIdx.x * kRowsPerBlock + row, k = blockIdx.y;
  double2* a = sh + row * 2 * M;
  double2* b = a + M;
  const bool ok = j <= g.ny;
  double* ar = reinterpret_cast<double*>(a);
  for (int i = t; i < 2 * M; i += kRowThreads) ar[i] = ok ? smp(g, i, j, k) : 0.0;
  __syncthreads();
  double2* z = fft_lds(a, b, t, kRowThreads, p, -1);
  if (!ok) return;
  // X_k = E_k + e^{-2 pi i k/nx} O_k with E, O from Z_k and conj(Z_{M-k})
  for (int q = t; q <= M; q += kRowThreads) {
    double2 zk = z[q % M], zc = conjd(z[(M - q) % M]);
    double2 e = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
    double2 dd = csub(zk, zc);
    double2 o = make_double2(0.5 * dd.y, -0.5 * dd.x);  // -i/2 (zk - zc)
    double2 x = cadd(e, cmul(twl[q], o));
    long long off = 2 * g.spec_at(q, j, k);
    phit[off] = x.x;
    phit[off + 1] = x.y;
  }
}

// z pass in place over PHIT columns (I, j) for I = blockIdx.x*C + c,
// j = 1 + blockIdx.y.  Forward divides by nx*nz (p2s_fftw2d.f:101).
constexpr int kCols = 4;
__global__ __launch_bounds__(256) void k_cfft_z(Geom g, double* phit, FftPlan p, int sign,
                                                int scale) {
  extern __shared__ double2 sh[];
  const int N = p.n;  // nz
  const int nh1 = g.nx / 2 + 1;
  const int I0 = blockIdx.x * kCols, j = 1 + blockIdx.y;
  // load: element e = c + C*k
  for (int e = threadIdx.x; e < kCols * N; e += blockDim.x) {
    int c = e % kCols, k = e / kCols;
    int I = I0 + c;
    double2 v = make_double2(0.0, 0.0);
    if (I < nh1) {
      long long off = 2 * g.spec_at(I, j, k);
      v = make_double2(phit[off], phit[off + 1]);
    }
    sh[c * 2 * N + k] = v;
  }
  __syncthreads();
  const int c = threadIdx.x / 64, t = threadIdx.x % 64;
  double2* res = fft_lds(sh + c * 2 * N, sh + c * 2 * N + N, t, 64, p, sign);
  // each column's result buffer is the same half for every column
  const int half = (res == sh + c * 2 * N) ? 0 : N;
  __syncthreads();
  const double sc = (double)(g.nx * g.nz);
  for (int e = threadIdx.x; e < kCols * N; e += blockDim.x) {
    int cc = e % kCols, k = e / kCols;
    int I = I0 + cc;
    if (I >= nh1) continue;
    double2 v = sh[cc * 2 * N + half + k];
    if (scale) v = make_double2(v.x / sc, v.y / sc);
    long long off = 2 * g.spec_at(I, j, k);
    phit[off] = v.x;
    phit[off + 1] = v.y;
  }
}

// inverse x pass (c2r) of row (j, k), then FP(0..nx-1, j, k), the wrap
// FP(nx, j, k) = FP(0, j, k) and, from row k = 0, the plane k = nz
// (p2s_fftw2d.f:205-216).
__global__ __launch_bounds__(kRowsPerBlock* kRowThreads) void k_irfft_x(Geom g,
                                                                         const double* phit,
                                                                         double* fp, FftPlan p,
                                                                         const double2* twl) {
  extern __shared__ double2 sh[];
  const int M = p.n;
  const int row = threadIdx.x / kRowThreads, t = threadIdx.x % kRowThreads;
  const int j = 1 + blockIdx.x * kRowsPerBlock + row, k = blockIdx.y;
  double2* a = sh + row * 2 * M;
  double2* b = a + M;
  const bool ok = j <= g.ny;
  // Z_k = (X_k + conj X_{M-k}) + i e^{2 pi i k/nx} (X_k - conj X_{M-k}),
  // with the imaginary parts of X_0 and X_M dropped as a c2r transform does
  for (int q = t; q < M; q += kRowThreads) {
    double2 xk = make_double2(0.0, 0.0), xc = make_double2(0.0, 0.0);
    if (ok) {
      long long o1 = 2 * g.spec_at(q, j, k), o2 = 2 * g.spec_at(M - q, j, k);
      xk = make_double2(phit[o1], q == 0 ? 0.0 : phit[o1 + 1]);
      xc = make_double2(phit[o2], (q == 0) ? 0.0 : -phit[o2 + 1]);
    }
    double2 sum = cadd(xk, xc), dif = csub(xk, xc);
    double2 r = cmul(conjd(twl[q]), dif);
    a[q] = make_double2(sum.x - r.y, sum.y + r.x);
  }
  __syncthreads();
  double2* z = fft_lds(a, b, t, kRowThreads, p, +1);
  if (!ok) return;
  const double* zr = reinterpret_cast<const double*>(z);
  for (int i = t; i < 2 * M; i += kRowThreads) {
    double x = zr[i];
    fp[g.at(i, j, k)] = x;
    if (i == 0) fp[g.at(g.nx, j, k)] = x;
    if (k == 0) {
      fp[g.at(i, j, g.nz)] = x;
      if (i == 0) fp[g.at(g.nx, j, g.nz)] = x;
    }
  }
}

static int plans_for(const Geom& g, FftPlan& px, FftPlan& pz, const double2*& twl) {
  if (g.nx % 2 != 0) return fail(VK_ERR_BAD_SIZE, "FFT: nx must be even");
  if (!fft_plan(g.nx / 2, px) || !fft_plan(g.nz, pz))
    return fail(VK_ERR_BAD_SIZE, "FFT: nx/2 and nz must be products of 2 and 3");
  twl = twiddle_table(g.nx);
  if (!twl) return fail(VK_ERR_HIP, "FFT: twiddle table");
  if (g.nz > 2048 || g.nx > 4096) return fail(VK_ERR_BAD_SIZE, "FFT: size above the LDS plan");
  return VK_OK;
}

template <typename Sample>
static int launch_forward(const Geom& g, Sample smp, double* phit, const char* xname,
                          hipStream_t s) {
  FftPlan px, pz;
  const double2* twl;
  int rc = plans_for(g, px, pz, twl);
  if (rc) return rc;
  size_t shx = sizeof(double2) * 2 * (g.nx / 2) * kRowsPerBlock;
  {
    ProfScope ps(xname, s);
    k_rfft_x<Sample><<<dim3((g.ny + kRowsPerBlock - 1) / kRowsPerBlock, g.nz),
                       kRowsPerBlock * kRowThreads, shx, s>>>(g, smp, phit, px, twl);
  }
  size_t shz = sizeof(double2) * 2 * g.nz * kCols;
  ProfScope ps("cfft_z_fwd", s);
  k_cfft_z<<<dim3((g.nx / 2 + 1 + kCols - 1) / kCols, g.ny), 256, shz, s>>>(g, phit, pz, -1, 1);
  return VK_OK;
}

int launch_p2s(const Geom& g, const double* fp, double* phit, hipStream_t s) {
  return launch_forward(g, SampleField{fp}, phit, "rfft_x", s);
}

int launch_p2s_div(const Geom& g, const Metrics& m, const double* u, const double* v,
                   const double* w, double dt, double* phit, hipStream_t s) {
  return launch_forward(g, SampleDiv{u, v, w, m.dx, m.dy, m.dz, dt}, phit, "rfft_x_div", s);
}

// S2P from a spectrum the call may overwrite (the z pass runs in place)
int launch_s2p_inplace(const Geom& g, double* phit, double* fp, hipStream_t s) {
  FftPlan px, pz;
  const double2* twl;
  int rc = plans_for(g, px, pz, twl);
  if (rc) return rc;
  size_t shz = sizeof(double2) * 2 * g.nz * kCols;
  {
    ProfScope ps("cfft_z_inv", s);
    k_cfft_z<<<dim3((g.nx / 2 + 1 + kCols - 1) / kCols, g.ny), 256, shz, s>>>(g, phit, pz, +1, 0);
  }
  size_t shx = sizeof(double2) * 2 * (g.nx / 2) * kRowsPerBlock;
  ProfScope ps("irfft_x", s);
  k_irfft_x<<<dim3((g.ny + kRowsPerBlock - 1) / kRowsPerBlock, g.nz),
              kRowsPerBlock * kRowThreads, shx, s>>>(g, phit, fp, px, twl);
  return VK_OK;
}

}  // namespace vk

using namespace vk;

extern "C" {

int vk_p2s(const double* fp, double* fpt, const int* nx, const int* ny, const int* nz) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  VK_RESOLVE(dfp, const double, fp, g.cells() * sizeof(double));
  VK_RESOLVE(dft, double, fpt, 2 * g.spec_cells() * sizeof(double));
  int rc = launch_p2s(g, dfp, dft, rt().stream);
  if (rc) return rc;
  VK_LAUNCH_CHECK("p2s");
  return VK_OK;
}

int vk_s2p(const double* fpt, double* fp, const int* nx, const int* ny, const int* nz) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  size_t sb = 2 * g.spec_cells() * sizeof(double);
  VK_RESOLVE(dft, const double, fpt, sb);
  VK_RESOLVE(dfp, double, fp, g.cells() * sizeof(double));
  double* work = static_cast<double*>(workspace(kWorkSpec2, sb));
  if (!work) return VK_ERR_HIP;
  int rc = hip_check(hipMemcpyAsync(work, dft, sb, hipMemcpyDeviceToDevice, rt().stream),
                     "s2p copy");
  if (rc) return rc;
  rc = launch_s2p_inplace(g, work, dfp, rt().stream);
  if (rc) return rc;
  VK_LAUNCH_CHECK("s2p");
  return VK_OK;
}

}  // extern "C"
