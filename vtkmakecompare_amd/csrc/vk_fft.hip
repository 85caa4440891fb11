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

// complex type and table of real type R
template <typename R>
struct Cx;
template <>
struct Cx<double> {
  using T = double2;
};
template <>
struct Cx<float> {
  using T = float2;
};

template <typename R>
struct FftPlan {
  using C = typename Cx<R>::T;
  int n = 0;
  int nst = 0;
  int r[20] = {0};
  const C* tw = nullptr;  // e^{-2 pi i m / n}, m = 0..n-1
};

namespace {
template <typename R>
std::map<int, FftPlan<R>>& plans() {
  static std::map<int, FftPlan<R>> p;
  return p;
}
template <typename R>
std::map<int, const typename Cx<R>::T*>& tables() {
  static std::map<int, const typename Cx<R>::T*> t;
  return t;
}
// vk_finalize hook: the twiddle tables are device memory; the plans point
// into them
void release_fft_tables() {
  for (auto& kv : tables<double>()) (void)dev_free(const_cast<double2*>(kv.second));
  for (auto& kv : tables<float>()) (void)dev_free(const_cast<float2*>(kv.second));
  tables<double>().clear();
  tables<float>().clear();
  plans<double>().clear();
  plans<float>().clear();
}
}  // namespace

// e^{-2 pi i m / n} for m = 0..n-1 in R, rounded once from extended
// precision, built once per (n, R).
template <typename R>
const typename Cx<R>::T* twiddle_table(int n) {
  using C = typename Cx<R>::T;
  on_finalize(release_fft_tables);
  auto it = tables<R>().find(n);
  if (it != tables<R>().end()) return it->second;
  std::vector<C> h(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; m++) {
    long double ang = -2.0L * pi * (long double)m / (long double)n;
    h[m].x = (R)cosl(ang);
    h[m].y = (R)sinl(ang);
  }
  C* d = nullptr;
  if (dev_malloc(reinterpret_cast<void**>(&d), sizeof(C) * n, "twiddle table") != hipSuccess)
    return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(C) * n, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  tables<R>()[n] = d;
  return d;
}

// Factor n into radix 4, 3, 2 passes; false if another prime divides n.
template <typename R>
static bool fft_plan(int n, FftPlan<R>& p) {
  auto it = plans<R>().find(n);
  if (it != plans<R>().end()) {
    p = it->second;
    return true;
  }
  FftPlan<R> q;
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
  q.tw = twiddle_table<R>(n);
  if (!q.tw) return false;
  plans<R>()[n] = q;
  p = q;
  return true;
}

template <typename C>
__device__ inline C mkc(decltype(C().x) x, decltype(C().x) y) {
  C c;
  c.x = x;
  c.y = y;
  return c;
}
template <typename C>
__device__ inline C cmul(C a, C b) {
  return mkc<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C>
__device__ inline C cadd(C a, C b) {
  return mkc<C>(a.x + b.x, a.y + b.y);
}
template <typename C>
__device__ inline C csub(C a, C b) {
  return mkc<C>(a.x - b.x, a.y - b.y);
}
template <typename C>
__device__ inline C conjc(C a) {
  return mkc<C>(a.x, -a.y);
}
template <typename C>
__device__ inline C cscale(decltype(C().x) s, C a) {
  return mkc<C>(s * a.x, s * a.y);
}
// multiply by sign * i
template <typename C>
__device__ inline C mul_si(C a, int sign) {
  return sign < 0 ? mkc<C>(a.y, -a.x) : mkc<C>(-a.y, a.x);
}

// Stockham autosort FFT of length p.n in LDS, threads t = 0..T-1 of one
// group.  sign = -1 forward, +1 backward, unnormalised.  Returns the
// buffer holding the result (a or b).  All threads of the block must call
// it together (it contains block barriers).
// One Stockham butterfly in place on v0..v3 (radix Rd = 4, 3 or 2; v3, or
// v2 and v3, unused below radix 4): the inputs times w^r for w =
// tw[kk twstep] (conjugated for sign > 0; one table read, the higher
// powers by products; kk twstep < n / Rd, so the index never wraps), then
// the radix-Rd DFT.  Scalars rather than arrays, so that nothing is
// addressed through memory.
template <typename R>
__device__ __forceinline__ void stockham_bfly(typename Cx<R>::T& v0, typename Cx<R>::T& v1,
                                              typename Cx<R>::T& v2, typename Cx<R>::T& v3,
                                              int Rd, int kk, int twstep,
                                              const typename Cx<R>::T* tw, int sign) {
  using C = typename Cx<R>::T;
  if (kk > 0) {
    C w1 = tw[kk * twstep];
    if (sign > 0) w1 = conjc(w1);
    v1 = cmul(v1, w1);
    if (Rd > 2) {
      const C w2 = cmul(w1, w1);
      v2 = cmul(v2, w2);
      if (Rd > 3) v3 = cmul(v3, cmul(w2, w1));
    }
  }
  if (Rd == 4) {
    const C s02 = cadd(v0, v2), d02 = csub(v0, v2);
    const C s13 = cadd(v1, v3), d13 = mul_si(csub(v1, v3), sign);
    v0 = cadd(s02, s13);
    v2 = csub(s02, s13);
    v1 = cadd(d02, d13);
    v3 = csub(d02, d13);
  } else if (Rd == 3) {
    const R h = (R)0.86602540378443864676;  // sqrt(3)/2
    const C sm = cadd(v1, v2), df = csub(v1, v2);
    const C a0 = v0;
    v0 = cadd(a0, sm);
    const C tt = mkc<C>(a0.x - (R)0.5 * sm.x, a0.y - (R)0.5 * sm.y);
    const C rot = mul_si(cscale(h, df), sign);
    v1 = cadd(tt, rot);
    v2 = csub(tt, rot);
  } else {
    const C a0 = v0;
    v0 = cadd(a0, v1);
    v1 = csub(a0, v1);
  }
}

// Passes s_begin .. s_end-1 (all by default) from a into the pair (a, b);
// returns the buffer that holds the result.  A barrier follows each pass.
template <typename R>
__device__ typename Cx<R>::T* fft_lds(typename Cx<R>::T* a, typename Cx<R>::T* b, int t, int T,
                                      const FftPlan<R>& p, const typename Cx<R>::T* tw, int sign,
                                      int s_begin = 0, int s_end = -1) {
  using C = typename Cx<R>::T;
  const int n = p.n;
  if (s_end < 0) s_end = p.nst;
  int ns = 1;
  for (int s = 0; s < s_begin; s++) ns *= p.r[s];
  C* src = a;
  C* dst = b;
  for (int s = s_begin; s < s_end; s++) {
    const int Rd = p.r[s];
    const int nb = n / Rd;
    const int twstep = n / (ns * Rd);
    for (int j = t; j < nb; j += T) {
      const int jq = j / ns, kk = j - jq * ns;
      C v0 = src[j], v1 = src[j + nb], v2 = v0, v3 = v0;
      if (Rd > 2) v2 = src[j + 2 * nb];
      if (Rd > 3) v3 = src[j + 3 * nb];
      stockham_bfly<R>(v0, v1, v2, v3, Rd, kk, twstep, tw, sign);
      const int base = jq * ns * Rd + kk;
      dst[base] = v0;
      dst[base + ns] = v1;
      if (Rd > 2) dst[base + 2 * ns] = v2;
      if (Rd > 3) dst[base + 3 * ns] = v3;
    }
    __syncthreads();
    C* tmp = src;
    src = dst;
    dst = tmp;
    ns *= Rd;
  }
  return src;
}

// Which sample feeds the forward x pass: a stored field, the Poisson
// right-hand side div(u_temp)/dt computed on the fly (mkp.f:88-121, where
// index 0 of x and z is the periodic image of nx and nz), or the residual
// of a Poisson solution against that right-hand side.  Samples are
// computed in FP64 whatever the transform's precision.  post() runs once
// per interior sample after every load of its chunk has been issued (a
// store among the loads would fence the ones after it).
struct SampleField {
  static constexpr int kChunk = 0;  // samples per chunk, 0: by precision
  const double* fp;
  __device__ double operator()(const Geom& g, int i, int j, int k) const {
    return fp[g.at(i, j, k)];
  }
  __device__ void post(const Geom&, int, int, int, double) const {}
};
// (mkp.f:88-99 divides by the metrics and by dt; multiplying by their
// reciprocals agrees to rounding)
struct SampleDiv {
  static constexpr int kChunk = 0;
  const double *u, *v, *w, *rdx, *rdy, *rdz;
  double rdt;
  __device__ double operator()(const Geom& g, int i, int j, int k) const {
    int ii = (i == 0) ? g.nx : i, kk = (k == 0) ? g.nz : k;
    long long o = g.at(ii, j, kk);
    return ((u[o] - u[o - 1]) * rdx[ii] + (v[o] - v[o - g.sy]) * rdy[j] +
            (w[o] - w[o - g.sz]) * rdz[kk]) *
           rdt;
  }
  __device__ void post(const Geom&, int, int, int, double) const {}
};
// div(u_temp)/dt, also stored in FP64 at its interior point (every
// interior point is sampled exactly once by the forward x pass)
struct SampleDivStore {
  static constexpr int kChunk = 0;
  SampleDiv div;
  double* out;
  __device__ double operator()(const Geom& g, int i, int j, int k) const {
    return div(g, i, j, k);
  }
  __device__ void post(const Geom& g, int i, int j, int k, double v) const {
    out[g.at((i == 0) ? g.nx : i, j, (k == 0) ? g.nz : k)] = v;
  }
};

// rhs - L phi for a float solution phi with filled halos and the stored
// FP64 right-hand side rhs = div(u_temp)/dt, L the discrete divergence of
// the discrete gradient that the projection applies (mkp.f:141-151,
// 209-217): the divergence the projected velocity would have, over dt.
// All FP64.
struct SampleResid {
  static constexpr int kChunk = 2;
  const double* rhs;
  const float* phi;
  const double *rdx12, *rdy12, *rdz12, *rdx, *rdy, *rdz;
  __device__ double operator()(const Geom& g, int i, int j, int k) const {
    int ii = (i == 0) ? g.nx : i, kk = (k == 0) ? g.nz : k;
    long long o = g.at(ii, j, kk);
    const double c = phi[o];
    const double lx = (((double)phi[o + 1] - c) * rdx12[ii] -
                       (c - (double)phi[o - 1]) * rdx12[ii - 1]) * rdx[ii];
    const double ly = (((double)phi[o + g.sy] - c) * rdy12[j] -
                       (c - (double)phi[o - g.sy]) * rdy12[j - 1]) * rdy[j];
    const double lz = (((double)phi[o + g.sz] - c) * rdz12[kk] -
                       (c - (double)phi[o - g.sz]) * rdz12[kk - 1]) * rdz[kk];
    return rhs[o] - (lx + ly + lz);
  }
  __device__ void post(const Geom&, int, int, int, double) const {}
};

// f(i,j,k) / dy(j) with x and z index 0 read as the periodic image n
struct SampleWScaled {
  static constexpr int kChunk = 0;
  const double *f, *rdy;
  __device__ double operator()(const Geom& g, int i, int j, int k) const {
    int ii = (i == 0) ? g.nx : i, kk = (k == 0) ? g.nz : k;
    return f[g.at(ii, j, kk)] * rdy[j];
  }
  __device__ void post(const Geom&, int, int, int, double) const {}
};

constexpr int kRowsPerBlock = 4;
constexpr int kRowThreads = 64;

// Samples a row thread evaluates before it stores any, so that all their
// loads are in flight together: a round of loads per chunk rather than
// per sample (the samples are FP64 and take more registers in the FP64
// transform).
template <typename Sample, typename R>
constexpr int sample_chunk() {
  return Sample::kChunk > 0 ? Sample::kChunk : (sizeof(R) == 4 ? 6 : 3);
}
// The float transforms read their twiddle tables from LDS, staged once per
// block with the first loads: a pass then waits on no global load.  (The
// FP64 rows fill LDS to its occupancy limit already, and the FP64
// transforms are bound by HBM bandwidth, not by these waits.)
template <typename R>
constexpr bool twiddles_in_lds() {
  return sizeof(R) == 4;
}
// twiddle entries a thread of an x block stages at most (2 M + 1 entries,
// M = nx / 2 <= 512)
constexpr int kTwPerThread = (2 * 512 + 1 + 255) / 256;

// LDS bytes of an x block: four rows of 2M complex, and with the twiddles
// in LDS the M entries of the plan's table and the M + 1 of the split's
template <typename R>
inline size_t xrow_lds(int nx) {
  const size_t m = (size_t)nx / 2;
  return sizeof(typename Cx<R>::T) *
         (2 * m * kRowsPerBlock + (twiddles_in_lds<R>() ? 2 * m + 1 : 0));
}

// forward x pass: rows (j, k) = (1 + blockIdx.x*4 + row, blockIdx.y)
template <typename Sample, typename R>
__global__ __launch_bounds__(kRowsPerBlock* kRowThreads) void k_rfft_x(
    Geom g, Sample smp, R* phit, FftPlan<R> p, const typename Cx<R>::T* twl) {
  using C = typename Cx<R>::T;
  constexpr int NT = kRowsPerBlock * kRowThreads;
  constexpr int CH = sample_chunk<Sample, R>();
  constexpr bool TWL = twiddles_in_lds<R>();
  extern __shared__ unsigned char shraw[];
  C* sh = reinterpret_cast<C*>(shraw);
  const int M = p.n;  // nx/2
  const int row = threadIdx.x / kRowThreads, t = threadIdx.x % kRowThreads;
  int bx, by, bz;
  xcd_block(bx, by, bz);
  const int j = 1 + bx * kRowsPerBlock + row, k = by;
  C* a = sh + row * 2 * M;
  C* b = a + M;
  const bool ok = j <= g.ny;
  const int js = ok ? j : g.ny;  // rows past ny sample row ny and keep zeros
  R* ar = reinterpret_cast<R*>(a);
  // twiddles: entries 0..M-1 the plan's table, M..2M the split's
  C* tws = sh + kRowsPerBlock * 2 * M;
  C tv[kTwPerThread];
  if (TWL) {
#pragma unroll
    for (int u = 0; u < kTwPerThread; u++) {
      const int q = (int)threadIdx.x + u * NT;
      if (q <= 2 * M) tv[u] = q < M ? p.tw[q] : twl[q - M];
    }
  }
  for (int i0 = 0; i0 < 2 * M; i0 += CH * kRowThreads) {
    double v[CH];
#pragma unroll
    for (int c = 0; c < CH; c++) {
      const int i = i0 + c * kRowThreads + t;
      v[c] = smp(g, i < 2 * M ? i : 2 * M - 1, js, k);
    }
#pragma unroll
    for (int c = 0; c < CH; c++) {
      const int i = i0 + c * kRowThreads + t;
      if (i < 2 * M) {
        ar[i] = ok ? (R)v[c] : (R)0;
        if (ok) smp.post(g, i, j, k, v[c]);
      }
    }
  }
  if (TWL) {
#pragma unroll
    for (int u = 0; u < kTwPerThread; u++) {
      const int q = (int)threadIdx.x + u * NT;
      if (q <= 2 * M) tws[q] = tv[u];
    }
  }
  __syncthreads();
  const C* tw1 = p.tw;
  const C* tw2 = twl;
  if constexpr (TWL) {
    tw1 = tws;
    tw2 = tws + M;
  }
  C* z = fft_lds<R>(a, b, t, kRowThreads, p, tw1, -1);
  if (!ok) return;
  // X_k = E_k + e^{-2 pi i k/nx} O_k with E, O from Z_k and conj(Z_{M-k})
  for (int q = t; q <= M; q += kRowThreads) {
    C zk = z[q % M], zc = conjc(z[(M - q) % M]);
    C e = mkc<C>((R)0.5 * (zk.x + zc.x), (R)0.5 * (zk.y + zc.y));
    C dd = csub(zk, zc);
    C o = mkc<C>((R)0.5 * dd.y, (R)-0.5 * dd.x);  // -i/2 (zk - zc)
    C x = cadd(e, cmul(tw2[q], o));
    long long off = 2 * g.spec_at(q, j, k);
    phit[off] = x.x;
    phit[off + 1] = x.y;
  }
}

// (forward: p2s_fftw2d.f:101 divides by nx nz)
// The z pass on CC adjacent columns (I0..I0+CC-1, j) of the spectrum.  The
// lanes take the columns in turn (lane = c + CC t), so every global access
// of a wave covers CC adjacent complex numbers per k.  The first Stockham
// pass reads its inputs straight from the spectrum and the last writes its
// outputs straight back (scaled by 1/(nx nz) on the forward pass); the
// passes between ping-pong in LDS, one padded buffer pair per column.  The
// first pass's reads all precede, across a barrier, the last pass's
// writes, except with one pass, which is a single butterfly per column.
template <typename R, int CC>
__global__ __launch_bounds__(256) void k_cfft_z(Geom g, R* phit, FftPlan<R> p, int sign,
                                                int scale) {
  using C = typename Cx<R>::T;
  constexpr int TPC = 256 / CC;
  extern __shared__ unsigned char shraw[];
  C* sh = reinterpret_cast<C*>(shraw);
  const int N = p.n;  // nz
  const int nh1 = g.nx / 2 + 1;
  int bx, by, bz;
  xcd_block(bx, by, bz);
  const int c = threadIdx.x % CC, t = threadIdx.x / CC;
  const int I = bx * CC + c, jrow = 1 + by;
  const bool okc = I < nh1;
  const long long o0 = 2 * g.spec_at(okc ? I : 0, jrow, 0);
  const long long ks = 2 * g.spec_at(0, 0, 1);  // k stride of the spectrum
  const R rsc = scale ? (R)(1.0 / ((double)g.nx * (double)g.nz)) : (R)1;
  C* src = sh + c * (2 * N + 1);
  C* dst = src + N;
  int ns = 1;
  for (int s = 0; s < p.nst; s++) {
    const int Rd = p.r[s];
    const int nb = N / Rd;
    const int twstep = N / (ns * Rd);
    const bool first = s == 0, last = s == p.nst - 1;
    for (int jj = t; jj < nb; jj += TPC) {
      const int jq = jj / ns, kk = jj - jq * ns;
      C v0, v1, v2, v3;
      if (first) {
        const C z = mkc<C>((R)0, (R)0);
        const long long off = o0 + ks * jj, st = ks * nb;
        v0 = okc ? mkc<C>(phit[off], phit[off + 1]) : z;
        v1 = okc ? mkc<C>(phit[off + st], phit[off + st + 1]) : z;
        v2 = (okc && Rd > 2) ? mkc<C>(phit[off + 2 * st], phit[off + 2 * st + 1]) : z;
        v3 = (okc && Rd > 3) ? mkc<C>(phit[off + 3 * st], phit[off + 3 * st + 1]) : z;
      } else {
        v0 = src[jj];
        v1 = src[jj + nb];
        v2 = Rd > 2 ? src[jj + 2 * nb] : v0;
        v3 = Rd > 3 ? src[jj + 3 * nb] : v0;
      }
      stockham_bfly<R>(v0, v1, v2, v3, Rd, kk, twstep, p.tw, sign);
      const int base = jq * ns * Rd + kk;
      if (last) {
        if (okc) {
          const long long off = o0 + ks * base, st = ks * ns;
          phit[off] = v0.x * rsc;
          phit[off + 1] = v0.y * rsc;
          phit[off + st] = v1.x * rsc;
          phit[off + st + 1] = v1.y * rsc;
          if (Rd > 2) {
            phit[off + 2 * st] = v2.x * rsc;
            phit[off + 2 * st + 1] = v2.y * rsc;
          }
          if (Rd > 3) {
            phit[off + 3 * st] = v3.x * rsc;
            phit[off + 3 * st + 1] = v3.y * rsc;
          }
        }
      } else {
        dst[base] = v0;
        dst[base + ns] = v1;
        if (Rd > 2) dst[base + 2 * ns] = v2;
        if (Rd > 3) dst[base + 3 * ns] = v3;
      }
    }
    if (!last) {
      __syncthreads();
      C* tmp = src;
      src = dst;
      dst = tmp;
    }
    ns *= Rd;
  }
}

// The z pass with one LDS buffer.  A thread takes W adjacent columns (W = 2
// in float, so that every access of a thread moves 16 bytes as in double),
// a block CC such column groups (CC W adjacent columns of the spectrum,
// one 128-byte piece of a row per k where LDS allows it).  Row jj of
// column group c sits at LDS slot (sw(jj) CC + c), sw(jj) = jj ^ ((jj >> 2)
// & 3): the rows a wave reads or writes in one instruction -- consecutive
// rows, rows Rd apart (the first pass's outputs), or groups of 4 rows 16
// apart -- then fall on both halves of the banks evenly, so no access
// conflicts.  Each pass loads all its inputs into registers, waits at a
// barrier, computes and stores its outputs, and waits again; a thread takes
// at most VPT rows per pass (butterflies jj = t, t + TPC, ...).  The first
// pass reads the spectrum and the last writes it, as in k_cfft_z.
__device__ __forceinline__ int zsw(int jj) { return jj ^ ((jj >> 2) & 3); }
constexpr int kZ1Threads = 512;  // threads per block
constexpr int kZ1Vpt = 6;        // rows per thread and pass, at most
// longest column the kernel takes (2 columns per block: kZ1Vpt rows for
// each of 256 threads)
constexpr int kZ1MaxN = kZ1Vpt * kZ1Threads / 2;
// LDS rows of a column buffer (rounded up to 4 for the swizzle)
inline __host__ __device__ size_t z1_rows(int n) { return ((size_t)n + 3) / 4 * 4; }

// (aligned to its size, so that a thread's LDS access is one 16-byte word)
template <typename R, int W>
struct alignas(sizeof(typename Cx<R>::T) * W) ZCols {
  typename Cx<R>::T c[W];
};

template <typename R, int W, int CC, int TPC, int Rd>
__device__ __forceinline__ void cfft_z_pass(R* phit, ZCols<R, W>* sh, const FftPlan<R>& p,
                                            const typename Cx<R>::T* tw, int sign, int N, int ns, bool first, bool last,
                                            int c, int t, const bool* okc, long long o0,
                                            long long ks, R rsc) {
  using C = typename Cx<R>::T;
  constexpr int BPT = (kZ1Vpt + Rd - 1) / Rd;
  const int nb = N / Rd;
  const int twstep = N / (ns * Rd);
  ZCols<R, W> v[BPT][4];
  const C z = mkc<C>((R)0, (R)0);
#pragma unroll
  for (int b = 0; b < BPT; b++) {
    const int jj = t + b * TPC;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int w = 0; w < W; w++) v[b][m].c[w] = z;
    if (jj < nb) {
#pragma unroll
      for (int m = 0; m < Rd; m++) {
        const int row = jj + m * nb;
        if (first) {
          const long long off = o0 + ks * row;
#pragma unroll
          for (int w = 0; w < W; w++)
            if (okc[w]) v[b][m].c[w] = mkc<C>(phit[off + 2 * w], phit[off + 2 * w + 1]);
        } else {
          v[b][m] = sh[zsw(row) * CC + c];
        }
      }
    }
  }
  if (!first && !last) __syncthreads();  // every read of the buffer precedes a write
#pragma unroll
  for (int b = 0; b < BPT; b++) {
    const int jj = t + b * TPC;
    if (jj < nb) {
      const int jq = jj / ns, kk = jj - jq * ns;
#pragma unroll
      for (int w = 0; w < W; w++)
        stockham_bfly<R>(v[b][0].c[w], v[b][1].c[w], v[b][2].c[w], v[b][3].c[w], Rd, kk, twstep,
                         tw, sign);
      const int base = jq * ns * Rd + kk;
#pragma unroll
      for (int m = 0; m < Rd; m++) {
        const int row = base + m * ns;
        if (last) {
          const long long off = o0 + ks * row;
#pragma unroll
          for (int w = 0; w < W; w++)
            if (okc[w]) {
              phit[off + 2 * w] = v[b][m].c[w].x * rsc;
              phit[off + 2 * w + 1] = v[b][m].c[w].y * rsc;
            }
        } else {
          sh[zsw(row) * CC + c] = v[b][m];
        }
      }
    }
  }
  if (!last) __syncthreads();
}

template <typename R, int W, int CC>
__global__ __launch_bounds__(kZ1Threads) void k_cfft_z1(Geom g, R* phit, FftPlan<R> p, int sign,
                                                        int scale) {
  constexpr int TPC = kZ1Threads / CC;
  extern __shared__ unsigned char shraw[];
  ZCols<R, W>* sh = reinterpret_cast<ZCols<R, W>*>(shraw);
  const int N = p.n;  // nz
  const int nh1 = g.nx / 2 + 1;
  int bx, by, bz;
  xcd_block(bx, by, bz);
  const int c = threadIdx.x % CC, t = threadIdx.x / CC;
  const int I = (bx * CC + c) * W, jrow = 1 + by;
  bool okc[W];
#pragma unroll
  for (int w = 0; w < W; w++) okc[w] = I + w < nh1;
  const long long o0 = 2 * g.spec_at(okc[0] ? I : 0, jrow, 0);
  const long long ks = 2 * g.spec_at(0, 0, 1);
  const R rsc = scale ? (R)(1.0 / ((double)g.nx * (double)g.nz)) : (R)1;
  // float: the twiddle table (N entries) staged in LDS past the column
  // buffer, loaded with the first pass's reads (which need no twiddle) and
  // stored after it
  using C = typename Cx<R>::T;
  constexpr bool TWL = twiddles_in_lds<R>();
  constexpr int TWP = kZ1MaxN / kZ1Threads;
  C* tws = reinterpret_cast<C*>(shraw + z1_rows(N) * CC * sizeof(ZCols<R, W>));
  C tv[TWP];
  if (TWL) {
#pragma unroll
    for (int u = 0; u < TWP; u++) {
      const int q = (int)threadIdx.x + u * kZ1Threads;
      if (q < N) tv[u] = p.tw[q];
    }
  }
  // the first pass reads no twiddle (ns = 1), so the plan may point at the
  // LDS copy from the start
  const C* tw1 = p.tw;
  if constexpr (TWL) tw1 = tws;
  int ns = 1;
  for (int s = 0; s < p.nst; s++) {
    const bool first = s == 0, last = s == p.nst - 1;
    switch (p.r[s]) {
      case 4:
        cfft_z_pass<R, W, CC, TPC, 4>(phit, sh, p, tw1, sign, N, ns, first, last, c, t, okc, o0, ks,
                                      rsc);
        break;
      case 3:
        cfft_z_pass<R, W, CC, TPC, 3>(phit, sh, p, tw1, sign, N, ns, first, last, c, t, okc, o0, ks,
                                      rsc);
        break;
      default:
        cfft_z_pass<R, W, CC, TPC, 2>(phit, sh, p, tw1, sign, N, ns, first, last, c, t, okc, o0, ks,
                                      rsc);
        break;
    }
    if (TWL && first) {
#pragma unroll
      for (int u = 0; u < TWP; u++) {
        const int q = (int)threadIdx.x + u * kZ1Threads;
        if (q < N) tws[q] = tv[u];
      }
      __syncthreads();
    }
    ns *= p.r[s];
  }
}

// k_cfft_z1 with the widest group of at most 128 bytes per k whose buffer
// (rows rounded up to 4 for the swizzle) fits in 64 KB and whose threads
// take at most kZ1Vpt rows per pass; false when none does (lengths above
// 3072 / 2 columns, or of one pass).
template <typename R>
static bool launch_cfft_z1(const Geom& g, R* phit, const FftPlan<R>& pz, int sign, int scale,
                           hipStream_t s) {
  constexpr int W = sizeof(R) == 4 ? 2 : 1;
  using E = ZCols<R, W>;
  const int N = g.nz;
  const int nh1 = g.nx / 2 + 1;
  const size_t rows = z1_rows(N);
  if (pz.nst < 2) return false;
  for (int cc = (int)(128 / sizeof(E)); cc >= 2; cc /= 2) {
    const size_t lds =
        rows * cc * sizeof(E) + (twiddles_in_lds<R>() ? N * sizeof(typename Cx<R>::T) : 0);
    const int tpc = kZ1Threads / cc;
    if (lds > 65536 || (N + tpc - 1) / tpc > kZ1Vpt) continue;
    const dim3 grid((unsigned)((nh1 + cc * W - 1) / (cc * W)), (unsigned)g.ny);
    switch (cc) {
      case 8:
        k_cfft_z1<R, W, 8><<<grid, kZ1Threads, lds, s>>>(g, phit, pz, sign, scale);
        return true;
      case 4:
        k_cfft_z1<R, W, 4><<<grid, kZ1Threads, lds, s>>>(g, phit, pz, sign, scale);
        return true;
      default:
        k_cfft_z1<R, W, 2><<<grid, kZ1Threads, lds, s>>>(g, phit, pz, sign, scale);
        return true;
    }
  }
  return false;
}

// the z pass with the widest column group whose LDS fits in 64 KB
template <typename R>
static void launch_cfft_z(const Geom& g, R* phit, const FftPlan<R>& pz, int sign, int scale,
                          hipStream_t s) {
  using C = typename Cx<R>::T;
  if (launch_cfft_z1<R>(g, phit, pz, sign, scale, s)) return;
  const size_t per_col = sizeof(C) * (2 * (size_t)g.nz + 1);
  const int nh1 = g.nx / 2 + 1;
  if (per_col * 8 <= 65536) {
    k_cfft_z<R, 8><<<dim3((nh1 + 7) / 8, g.ny), 256, per_col * 8, s>>>(g, phit, pz, sign, scale);
  } else if (per_col * 4 <= 65536) {
    k_cfft_z<R, 4><<<dim3((nh1 + 3) / 4, g.ny), 256, per_col * 4, s>>>(g, phit, pz, sign, scale);
  } else if (per_col * 2 <= 65536) {
    k_cfft_z<R, 2><<<dim3((nh1 + 1) / 2, g.ny), 256, per_col * 2, s>>>(g, phit, pz, sign, scale);
  } else {
    k_cfft_z<R, 1><<<dim3(nh1, g.ny), 256, per_col, s>>>(g, phit, pz, sign, scale);
  }
}

// inverse x pass (c2r) of row (j, k), then FP(0..nx-1, j, k), the wrap
// FP(nx, j, k) = FP(0, j, k) and, from row k = 0, the plane k = nz
// (p2s_fftw2d.f:205-216).
template <typename R>
__global__ __launch_bounds__(kRowsPerBlock* kRowThreads) void k_irfft_x(
    Geom g, const R* phit, R* fp, FftPlan<R> p, const typename Cx<R>::T* twl) {
  using C = typename Cx<R>::T;
  constexpr int NT = kRowsPerBlock * kRowThreads;
  constexpr bool TWL = twiddles_in_lds<R>();
  extern __shared__ unsigned char shraw[];
  C* sh = reinterpret_cast<C*>(shraw);
  const int M = p.n;
  const int row = threadIdx.x / kRowThreads, t = threadIdx.x % kRowThreads;
  // the plan's twiddles (M entries) staged in LDS in float, loaded here
  // with the first pass's reads and stored before its barrier; the first
  // pass itself needs none
  C* tws = sh + kRowsPerBlock * 2 * M;
  C tv[kTwPerThread];
  if (TWL) {
#pragma unroll
    for (int u = 0; u < kTwPerThread; u++) {
      const int q = (int)threadIdx.x + u * NT;
      if (q < M) tv[u] = p.tw[q];
    }
  }
  const C* tw1 = p.tw;
  if constexpr (TWL) tw1 = tws;
  int bx, by, bz;
  xcd_block(bx, by, bz);
  const int j = 1 + bx * kRowsPerBlock + row, k = by;
  C* a = sh + row * 2 * M;
  C* b = a + M;
  const bool ok = j <= g.ny;
  // Z_k = (X_k + conj X_{M-k}) + i e^{2 pi i k/nx} (X_k - conj X_{M-k}),
  // with the imaginary parts of X_0 and X_M dropped as a c2r transform does;
  // the first pass reads its inputs this way straight from the spectrum
  // (the loads are unconditional, rows past ny reading row ny, so that the
  // inputs of all a thread's butterflies are in flight together)
  const int js = ok ? j : g.ny;
  auto zin = [&](int q) {
    const long long o1 = 2 * g.spec_at(q, js, k), o2 = 2 * g.spec_at(M - q, js, k);
    const R x1r = phit[o1], x1i = phit[o1 + 1], x2r = phit[o2], x2i = phit[o2 + 1];
    const C zero = mkc<C>((R)0, (R)0);
    const C xk = ok ? mkc<C>(x1r, q == 0 ? (R)0 : x1i) : zero;
    const C xc = ok ? mkc<C>(x2r, q == 0 ? (R)0 : -x2i) : zero;
    C sum = cadd(xk, xc), dif = csub(xk, xc);
    C r = cmul(conjc(twl[q]), dif);
    return mkc<C>(sum.x - r.y, sum.y + r.x);
  };
  // the last pass writes FP(0..nx-1, j, k) straight to the field, with the
  // wrap FP(nx, j, k) = FP(0, j, k) and, from row k = 0, the plane k = nz
  // (p2s_fftw2d.f:205-216)
  auto out = [&](int m, C v) {
    const int i = 2 * m;
    fp[g.at(i, j, k)] = v.x;
    fp[g.at(i + 1, j, k)] = v.y;
    if (i == 0) fp[g.at(g.nx, j, k)] = v.x;
    if (k == 0) {
      fp[g.at(i, j, g.nz)] = v.x;
      fp[g.at(i + 1, j, g.nz)] = v.y;
      if (i == 0) fp[g.at(g.nx, j, g.nz)] = v.x;
    }
  };
  const int last = p.nst - 1;
  int ns = 1;
  for (int s = 0; s < last; s++) ns *= p.r[s];
  if (last == 0) {
    // one pass (nx <= 8): one butterfly per row, read and written by one thread
    const int Rd = p.r[0], nb = M / Rd;
    for (int jj = t; jj < nb; jj += kRowThreads) {
      C v0 = zin(jj), v1 = zin(jj + nb), v2 = v0, v3 = v0;
      if (Rd > 2) v2 = zin(jj + 2 * nb);
      if (Rd > 3) v3 = zin(jj + 3 * nb);
      stockham_bfly<R>(v0, v1, v2, v3, Rd, 0, 0, p.tw, +1);
      if (ok) {
        out(jj * Rd, v0);
        out(jj * Rd + 1, v1);
        if (Rd > 2) out(jj * Rd + 2, v2);
        if (Rd > 3) out(jj * Rd + 3, v3);
      }
    }
    return;
  }
  {
    const int Rd = p.r[0], nb = M / Rd;
    for (int jj = t; jj < nb; jj += kRowThreads) {
      C v0 = zin(jj), v1 = zin(jj + nb), v2 = v0, v3 = v0;
      if (Rd > 2) v2 = zin(jj + 2 * nb);
      if (Rd > 3) v3 = zin(jj + 3 * nb);
      stockham_bfly<R>(v0, v1, v2, v3, Rd, 0, 0, p.tw, +1);
      const int base = jj * Rd;
      b[base] = v0;
      b[base + 1] = v1;
      if (Rd > 2) b[base + 2] = v2;
      if (Rd > 3) b[base + 3] = v3;
    }
    if (TWL) {
#pragma unroll
      for (int u = 0; u < kTwPerThread; u++) {
        const int q = (int)threadIdx.x + u * NT;
        if (q < M) tws[q] = tv[u];
      }
    }
    __syncthreads();
  }
  const C* z = fft_lds<R>(b, a, t, kRowThreads, p, tw1, +1, 1, last);
  {
    const int Rd = p.r[last], nb = M / Rd;
    const int twstep = M / (ns * Rd);
    for (int jj = t; jj < nb; jj += kRowThreads) {
      const int jq = jj / ns, kk = jj - jq * ns;
      C v0 = z[jj], v1 = z[jj + nb], v2 = v0, v3 = v0;
      if (Rd > 2) v2 = z[jj + 2 * nb];
      if (Rd > 3) v3 = z[jj + 3 * nb];
      stockham_bfly<R>(v0, v1, v2, v3, Rd, kk, twstep, tw1, +1);
      const int base = jq * ns * Rd + kk;
      if (ok) {
        out(base, v0);
        out(base + ns, v1);
        if (Rd > 2) out(base + 2 * ns, v2);
        if (Rd > 3) out(base + 3 * ns, v3);
      }
    }
  }
}

// The sizes the spectral solve takes: nx even, nx/2 and nz products of 2
// and 3, one FP64 column (2 nz + 1 complex) and four FP64 x rows (2 nx complex
// each) within 64 KB of LDS.  Checked before a step launches anything, so
// a refused size leaves every field as it was.
int fft_sizes_ok(const Geom& g) {
  auto smooth = [](int n) {
    if (n < 1) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
  };
  if (g.nx % 2 != 0) return fail(VK_ERR_BAD_SIZE, "FFT: nx must be even");
  if (!smooth(g.nx / 2) || !smooth(g.nz))
    return fail(VK_ERR_BAD_SIZE, "FFT: nx/2 and nz must be products of 2 and 3");
  if ((2 * (size_t)g.nz + 1) * 16 > 65536 || 4 * 2 * (size_t)(g.nx / 2) * 16 > 65536)
    return fail(VK_ERR_BAD_SIZE, "FFT: size above the LDS plan");
  return VK_OK;
}

template <typename R>
static int plans_for(const Geom& g, FftPlan<R>& px, FftPlan<R>& pz,
                     const typename Cx<R>::T*& twl) {
  if (g.nx % 2 != 0) return fail(VK_ERR_BAD_SIZE, "FFT: nx must be even");
  if (!fft_plan<R>(g.nx / 2, px) || !fft_plan<R>(g.nz, pz))
    return fail(VK_ERR_BAD_SIZE, "FFT: nx/2 and nz must be products of 2 and 3");
  twl = twiddle_table<R>(g.nx);
  if (!twl) return fail(VK_ERR_HIP, "FFT: twiddle table");
  // one column (2 nz complex) and four x rows (2 nx complex each) in 64 KB
  if ((2 * (size_t)g.nz + 1) * sizeof(typename Cx<R>::T) > 65536 ||
      4 * 2 * (size_t)(g.nx / 2) * sizeof(typename Cx<R>::T) > 65536)
    return fail(VK_ERR_BAD_SIZE, "FFT: size above the LDS plan");
  return VK_OK;
}

template <typename R, typename Sample>
static int launch_forward(const Geom& g, Sample smp, R* phit, const char* xname,
                          const char* zname, hipStream_t s) {
  using C = typename Cx<R>::T;
  FftPlan<R> px, pz;
  const C* twl;
  int rc = plans_for<R>(g, px, pz, twl);
  if (rc) return rc;
  const size_t shx = xrow_lds<R>(g.nx);
  {
    ProfScope ps(xname, s);
    k_rfft_x<Sample, R><<<dim3((g.ny + kRowsPerBlock - 1) / kRowsPerBlock, g.nz),
                          kRowsPerBlock * kRowThreads, shx, s>>>(g, smp, phit, px, twl);
  }
  ProfScope ps(zname, s);
  launch_cfft_z<R>(g, phit, pz, -1, 1, s);
  return VK_OK;
}

int launch_p2s(const Geom& g, const double* fp, double* phit, hipStream_t s) {
  return launch_forward<double>(g, SampleField{fp}, phit, "rfft_x", "cfft_z_fwd", s);
}

template <typename R>
int launch_p2s_div(const Geom& g, const RMetrics& r, const double* u, const double* v,
                   const double* w, double dt, R* phit, hipStream_t s) {
  return launch_forward<R>(g, SampleDiv{u, v, w, r.rdx, r.rdy, r.rdz, 1.0 / dt}, phit,
                           "rfft_x_div",
                           "cfft_z_fwd", s);
}
template int launch_p2s_div<double>(const Geom&, const RMetrics&, const double*, const double*,
                                    const double*, double, double*, hipStream_t);

int launch_p2s_div_f32(const Geom& g, const RMetrics& r, const double* u, const double* v,
                       const double* w, double dt, float* phit, hipStream_t s, const char* xname,
                       const char* zname) {
  return launch_forward<float>(g, SampleDiv{u, v, w, r.rdx, r.rdy, r.rdz, 1.0 / dt}, phit, xname,
                               zname, s);
}

int launch_p2s_div_store(const Geom& g, const RMetrics& r, const double* u, const double* v,
                         const double* w, double dt, double* rhs, float* phit, hipStream_t s) {
  return launch_forward<float>(
      g, SampleDivStore{SampleDiv{u, v, w, r.rdx, r.rdy, r.rdz, 1.0 / dt}, rhs}, phit,
      "rfft_x_div", "cfft_z_fwd", s);
}

int launch_p2s_resid(const Geom& g, const RMetrics& r, const double* rhs, const float* phi,
                     float* phit, hipStream_t s) {
  SampleResid smp{rhs, phi, r.rdx12, r.rdy12, r.rdz12, r.rdx, r.rdy, r.rdz};
  return launch_forward<float>(g, smp, phit, "rfft_x_resid", "cfft_z_fwd2", s);
}

int launch_p2s_wscaled(const Geom& g, const double* f, const double* rdy, double* phit,
                       hipStream_t s) {
  return launch_forward<double>(g, SampleWScaled{f, rdy}, phit, "rfft_x_adj", "cfft_z_adj", s);
}

// S2P from a spectrum the call may overwrite (the z pass runs in place)
template <typename R>
int launch_s2p_inplace(const Geom& g, R* phit, R* fp, hipStream_t s, const char* zname,
                       const char* xname) {
  using C = typename Cx<R>::T;
  FftPlan<R> px, pz;
  const C* twl;
  int rc = plans_for<R>(g, px, pz, twl);
  if (rc) return rc;
  {
    ProfScope ps(zname, s);
    launch_cfft_z<R>(g, phit, pz, +1, 0, s);
  }
  const size_t shx = xrow_lds<R>(g.nx);
  ProfScope ps(xname, s);
  k_irfft_x<R><<<dim3((g.ny + kRowsPerBlock - 1) / kRowsPerBlock, g.nz),
                 kRowsPerBlock * kRowThreads, shx, s>>>(g, phit, fp, px, twl);
  return VK_OK;
}
template int launch_s2p_inplace<double>(const Geom&, double*, double*, hipStream_t, const char*,
                                        const char*);
template int launch_s2p_inplace<float>(const Geom&, float*, float*, hipStream_t, const char*,
                                       const char*);

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
  rc = launch_s2p_inplace<double>(g, work, dfp, rt().stream, "cfft_z_inv", "irfft_x");
  if (rc) return rc;
  VK_LAUNCH_CHECK("s2p");
  return VK_OK;
}

}  // extern "C"
