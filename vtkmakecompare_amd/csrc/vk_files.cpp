// vk_files.cpp -- the reference's legacy VTK files and the library's own
// restart record.
//
// VTK: levelset.vtk (set_levelset.f:131-156), cellvalue1.vtk
// (set_levelset.f:171-198) and cellvalue.vtk (nse_dominik.f:208-235) are
// ASCII structured grids written with Fortran edit descriptors; they are
// reproduced byte for byte: 'DIMENSIONS ',3(1x,i3) (asterisks from 1000
// points on an axis), 'POINTS ',i9,' float', the cell centres as 3e15.7,
// the values as e15.7.
//
// Restart: READFIELD (nse_dominik.f:2769-2793) restores u..w_ta but not the
// AB2 history nor the step counter, so a restarted run takes an Euler step
// where the uninterrupted one takes an AB2 step.  The restart record here
// carries everything the fused step reads: the state, the history, u_ta,
// rhoCp, thc, all four cellvalue slots and l, in Fortran sequential
// unformatted records (4-byte markers, one record per array).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "vk_runtime.h"

namespace vk {
namespace {

// ---------------------------------------------------------------- formats

// Fortran Ew.d output of v (no scale factor): [-]0.d1..dd E+ee, the
// exponent without the E when its magnitude needs three digits; decimal
// rounding to nearest with ties to even on the exact binary value (as the
// Fortran runtime and printf both round); Inf/NaN as text; right-justified
// in w columns, asterisks when it does not fit.
void fortran_e(double v, int w, int d, std::string& out) {
  char body[64];
  int len = 0;
  if (std::isnan(v)) {
    len = std::snprintf(body, sizeof(body), "NaN");
  } else if (std::isinf(v)) {
    len = std::snprintf(body, sizeof(body), v < 0 ? "-Inf" : "Inf");
  } else {
    const bool neg = std::signbit(v);
    const double a = std::fabs(v);
    char digits[40];
    int e10 = 0;
    if (a == 0.0) {
      std::memset(digits, '0', (size_t)d);
      digits[d] = '\0';
    } else {
      char buf[64];
      std::snprintf(buf, sizeof(buf), "%.*e", d - 1, a);  // d.ddd...e+xx
      int q = 0;
      digits[q++] = buf[0];
      const char* p = buf + 2;
      while (*p && *p != 'e') digits[q++] = *p++;
      digits[q] = '\0';
      e10 = std::atoi(p + 1) + 1;
    }
    char ex[16];
    const int ae = e10 < 0 ? -e10 : e10;
    if (ae <= 99)
      std::snprintf(ex, sizeof(ex), "E%c%02d", e10 < 0 ? '-' : '+', ae);
    else
      std::snprintf(ex, sizeof(ex), "%c%03d", e10 < 0 ? '-' : '+', ae);
    len = std::snprintf(body, sizeof(body), "%s0.%s%s", neg ? "-" : "", digits, ex);
  }
  if (len > w) {
    out.append((size_t)w, '*');
    return;
  }
  out.append((size_t)(w - len), ' ');
  out.append(body, (size_t)len);
}

// Fortran Iw output: right-justified, asterisks when it does not fit
void fortran_i(long long v, int w, std::string& out) {
  char buf[32];
  const int len = std::snprintf(buf, sizeof(buf), "%lld", v);
  if (len > w) {
    out.append((size_t)w, '*');
    return;
  }
  out.append((size_t)(w - len), ' ');
  out.append(buf, (size_t)len);
}

// --------------------------------------------------------- caller arrays

// A caller's array of n doubles on the host: a registered host array is
// brought up to date from its device mirror, device memory is copied into
// a buffer, any other pointer is host memory.  Device copies go on the
// library stream, after the work queued there (every entry point joins its
// streams into it; a copy on the null stream would not wait for them, as
// the library's streams do not block on it).
struct HostIn {
  const double* p = nullptr;
  std::vector<double> buf;
};

int host_in(const double* a, size_t n, HostIn& h, const char* what) {
  if (!a) return fail(VK_ERR_ARGUMENT, std::string(what) + " is null");
  if (is_registered(a)) {
    // the registration must cover the n values read here
    if (!resolve(a, n * sizeof(double), what)) return last_status();
    int rc = vk_download(const_cast<double*>(a));
    if (rc) return rc;
    h.p = a;
    return VK_OK;
  }
  if (is_device(a)) {
    size_t ext = 0;
    if (device_extent(a, &ext) && ext < n * sizeof(double))
      return fail(VK_ERR_ARGUMENT, std::string(what) + ": device allocation too small");
    h.buf.resize(n);
    int rc = hip_check(hipMemcpyAsync(h.buf.data(), a, n * sizeof(double),
                                      hipMemcpyDeviceToHost, rt().stream),
                       what);
    if (!rc) rc = hip_check(hipStreamSynchronize(rt().stream), what);
    if (rc) return rc;
    h.p = h.buf.data();
    return VK_OK;
  }
  h.p = a;
  return VK_OK;
}

// Where a read of n doubles for the caller's array goes, and how it gets
// there afterwards (finish): a registered host array is filled and
// uploaded, device memory filled through a buffer.
struct HostOut {
  double* dst = nullptr;  // the caller's pointer
  double* p = nullptr;    // where to read into
  std::vector<double> buf;
  bool dev = false, reg = false;
  size_t n = 0;
  int finish(const char* what) {
    if (dev) {
      // on the library stream, after the work queued there, like host_in
      int rc = hip_check(
          hipMemcpyAsync(dst, p, n * sizeof(double), hipMemcpyHostToDevice, rt().stream), what);
      if (!rc) rc = hip_check(hipStreamSynchronize(rt().stream), what);
      return rc;
    }
    if (reg) return vk_upload(dst);
    return VK_OK;
  }
};

int host_out(double* a, size_t n, HostOut& h, const char* what) {
  if (!a) return fail(VK_ERR_ARGUMENT, std::string(what) + " is null");
  h.dst = a;
  h.n = n;
  if (is_registered(a)) {
    if (!resolve(a, n * sizeof(double), what)) return last_status();
    h.reg = true;
    h.p = a;
    return VK_OK;
  }
  if (is_device(a)) {
    size_t ext = 0;
    if (device_extent(a, &ext) && ext < n * sizeof(double))
      return fail(VK_ERR_ARGUMENT, std::string(what) + ": device allocation too small");
    h.dev = true;
    h.buf.resize(n);
    h.p = h.buf.data();
    return VK_OK;
  }
  h.p = a;
  return VK_OK;
}

// ----------------------------------------------------------------- VTK

int threads_for(long long work) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const long long t = std::min<long long>(std::min<long long>(hw, 16), 1 + work / 200000);
  return (int)std::max<long long>(1, t);
}

// The structured-grid file of set_levelset.f:131-156 / :171-198 and
// nse_dominik.f:208-235: value(i, j, k) is the point data, `scalars` its
// SCALARS line.
template <typename F>
int write_vtk(const char* path, int nx, int ny, int nz, const double* x, const double* y,
              const double* z, const char* scalars, F value) {
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return fail(VK_ERR_ARGUMENT, std::string("vtk: cannot create ") + path);
  const long long npts = (long long)nx * ny * nz;
  std::string head = "# vtk DataFile Version 3.8.1\n3D structure of vessel\nASCII \n"
                     "DATASET STRUCTURED_GRID\nDIMENSIONS ";
  for (int n : {nx, ny, nz}) {
    head += ' ';
    fortran_i(n, 3, head);
  }
  head += "\nPOINTS ";
  fortran_i(npts, 9, head);
  head += " float\n";
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  // the coordinates of one axis formatted once: x(i) y(j) z(k) per line
  std::vector<std::string> fx(nx), fy(ny), fz(nz);
  for (int i = 1; i <= nx; i++) fortran_e(x[i], 15, 7, fx[i - 1]);
  for (int j = 1; j <= ny; j++) fortran_e(y[j], 15, 7, fy[j - 1]);
  for (int k = 1; k <= nz; k++) fortran_e(z[k], 15, 7, fz[k - 1]);
  // planes formatted in parallel, written in order
  const int nt = threads_for(npts);
  const int chunk = std::max(1, nt * 2);
  std::vector<std::string> planes((size_t)chunk);
  auto run = [&](auto&& fmt_plane) {
    for (int k0 = 1; k0 <= nz && ok; k0 += chunk) {
      const int k1 = std::min(nz, k0 + chunk - 1);
      std::vector<std::thread> th;
      for (int t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
          for (int k = k0 + t; k <= k1; k += nt) {
            std::string& s = planes[(size_t)(k - k0)];
            s.clear();
            fmt_plane(k, s);
          }
        });
      for (auto& t : th) t.join();
      for (int k = k0; k <= k1 && ok; k++) {
        const std::string& s = planes[(size_t)(k - k0)];
        ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
      }
    }
  };
  run([&](int k, std::string& s) {
    s.reserve((size_t)nx * ny * 46);
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        s += fx[i - 1];
        s += fy[j - 1];
        s += fz[k - 1];
        s += '\n';
      }
  });
  std::string mid = "POINT_DATA ";
  fortran_i(npts, 9, mid);
  mid += '\n';
  mid += scalars;
  mid += "\nLOOKUP_TABLE default\n";
  ok = ok && std::fwrite(mid.data(), 1, mid.size(), f) == mid.size();
  run([&](int k, std::string& s) {
    s.reserve((size_t)nx * ny * 16);
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        fortran_e(value(i, j, k), 15, 7, s);
        s += '\n';
      }
  });
  if (std::fclose(f) != 0) ok = false;
  return ok ? VK_OK : fail(VK_ERR_ARGUMENT, std::string("vtk: write failed on ") + path);
}

int vtk_args(const int* nx, const int* ny, const int* nz, const double* x, const double* y,
             const double* z, HostIn& hx, HostIn& hy, HostIn& hz) {
  if (!nx || !ny || !nz || *nx < 1 || *ny < 1 || *nz < 1)
    return fail(VK_ERR_BAD_SIZE, "vtk: nx, ny, nz >= 1");
  int rc = host_in(x, (size_t)*nx + 2, hx, "vtk: x");
  if (!rc) rc = host_in(y, (size_t)*ny + 2, hy, "vtk: y");
  if (!rc) rc = host_in(z, (size_t)*nz + 2, hz, "vtk: z");
  return rc;
}

// ------------------------------------------------------------- restart

constexpr char kRestartMagic[16] = {'V', 'K', 'R', 'E', 'S', 'T', 'A', 'R',
                                    'T', ' ', '1', ' ', ' ', ' ', ' ', ' '};
constexpr int kRestartFields = 14;  // halo fields, then 4 cellvalue slots

bool put_record(std::FILE* f, const void* data, size_t bytes) {
  const int32_t len = (int32_t)bytes;
  return std::fwrite(&len, 4, 1, f) == 1 &&
         (bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes) &&
         std::fwrite(&len, 4, 1, f) == 1;
}

bool get_record(std::FILE* f, void* data, size_t bytes) {
  int32_t a = 0, b = 0;
  if (std::fread(&a, 4, 1, f) != 1 || a < 0 || (size_t)a != bytes) return false;
  if (bytes && std::fread(data, 1, bytes, f) != bytes) return false;
  return std::fread(&b, 4, 1, f) == 1 && b == a;
}

}  // namespace
}  // namespace vk

using namespace vk;

extern "C" {

int vk_vtk_levelset(const int* nx, const int* ny, const int* nz, const double* x,
                    const double* y, const double* z, const double* phi, const char* path) {
  HostIn hx, hy, hz, hp;
  int rc = vtk_args(nx, ny, nz, x, y, z, hx, hy, hz);
  if (rc) return rc;
  const size_t sy = (size_t)*nx + 2, sz = sy * ((size_t)*ny + 2);
  rc = host_in(phi, sz * ((size_t)*nz + 2), hp, "vtk: phi");
  if (rc) return rc;
  const double* p = hp.p;
  return write_vtk(path ? path : "levelset.vtk", *nx, *ny, *nz, hx.p, hy.p, hz.p,
                   "SCALARS phi float64",
                   [&](int i, int j, int k) { return p[(size_t)i + sy * j + sz * k]; });
}

int vk_vtk_cellvalue(const int* nx, const int* ny, const int* nz, const double* x,
                     const double* y, const double* z, const double* cellvalue,
                     const char* path) {
  HostIn hx, hy, hz, hc;
  int rc = vtk_args(nx, ny, nz, x, y, z, hx, hy, hz);
  if (rc) return rc;
  const size_t n = (size_t)*nx * *ny * *nz;
  // slot 1 only: the first nx*ny*nz values
  rc = host_in(cellvalue, n, hc, "vtk: cellvalue");
  if (rc) return rc;
  const double* c = hc.p;
  const size_t sy = (size_t)*nx, sz = sy * (size_t)*ny;
  return write_vtk(path ? path : "cellvalue.vtk", *nx, *ny, *nz, hx.p, hy.p, hz.p,
                   "SCALARS PHI float", [&](int i, int j, int k) {
                     return c[(size_t)(i - 1) + sy * (j - 1) + sz * (k - 1)];
                   });
}

int vk_write_restart(const char* path, const int* nx, const int* ny, const int* nz,
                     const int* l, const double* u, const double* v, const double* w,
                     const double* p, const double* t, const double* u_pre,
                     const double* v_pre, const double* w_pre, const double* t_pre,
                     const double* u_ta, const double* v_ta, const double* w_ta,
                     const double* rhocp, const double* thc, const double* cellvalue) {
  if (!path) return fail(VK_ERR_ARGUMENT, "write_restart: path is null");
  if (!nx || !ny || !nz || !l || *nx < 1 || *ny < 1 || *nz < 1)
    return fail(VK_ERR_BAD_SIZE, "write_restart: nx, ny, nz >= 1");
  const size_t nf = ((size_t)*nx + 2) * ((size_t)*ny + 2) * ((size_t)*nz + 2);
  const size_t ni = (size_t)*nx * *ny * *nz;
  if (nf * sizeof(double) > 2147483647u)
    return fail(VK_ERR_BAD_SIZE, "write_restart: a field exceeds one 2 GiB record");
  const double* fields[kRestartFields] = {u,     v,     w,     p,    t,    u_pre, v_pre,
                                          w_pre, t_pre, u_ta,  v_ta, w_ta, rhocp, thc};
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return fail(VK_ERR_ARGUMENT, std::string("write_restart: cannot create ") + path);
  int32_t hdr[8] = {*nx, *ny, *nz, *l, kRestartFields, 4, 8, 0};
  char head[16 + sizeof(hdr)];
  std::memcpy(head, kRestartMagic, 16);
  std::memcpy(head + 16, hdr, sizeof(hdr));
  bool ok = put_record(f, head, sizeof(head));
  int rc = VK_OK;
  for (int q = 0; q < kRestartFields && ok && !rc; q++) {
    HostIn h;
    rc = host_in(fields[q], nf, h, "write_restart: field");
    if (!rc) ok = put_record(f, h.p, nf * sizeof(double));
  }
  if (ok && !rc) {  // the four slots, one record each
    HostIn h;
    rc = host_in(cellvalue, 4 * ni, h, "write_restart: cellvalue");
    for (int s = 0; s < 4 && ok && !rc; s++) ok = put_record(f, h.p + s * ni, ni * sizeof(double));
  }
  if (std::fclose(f) != 0) ok = false;
  if (rc) return rc;
  return ok ? VK_OK : fail(VK_ERR_ARGUMENT, std::string("write_restart: write failed on ") + path);
}

int vk_read_restart(const char* path, const int* nx, const int* ny, const int* nz, int* l,
                    double* u, double* v, double* w, double* p, double* t, double* u_pre,
                    double* v_pre, double* w_pre, double* t_pre, double* u_ta, double* v_ta,
                    double* w_ta, double* rhocp, double* thc, double* cellvalue) {
  if (!path || !l) return fail(VK_ERR_ARGUMENT, "read_restart: path or l is null");
  if (!nx || !ny || !nz || *nx < 1 || *ny < 1 || *nz < 1)
    return fail(VK_ERR_BAD_SIZE, "read_restart: nx, ny, nz >= 1");
  const size_t nf = ((size_t)*nx + 2) * ((size_t)*ny + 2) * ((size_t)*nz + 2);
  const size_t ni = (size_t)*nx * *ny * *nz;
  double* fields[kRestartFields] = {u,     v,     w,     p,    t,    u_pre, v_pre,
                                    w_pre, t_pre, u_ta,  v_ta, w_ta, rhocp, thc};
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return fail(VK_ERR_ARGUMENT, std::string("read_restart: cannot open ") + path);
  int32_t hdr[8];
  char head[16 + sizeof(hdr)];
  if (!get_record(f, head, sizeof(head)) || std::memcmp(head, kRestartMagic, 16) != 0) {
    std::fclose(f);
    return fail(VK_ERR_ARGUMENT, std::string("read_restart: not a restart record: ") + path);
  }
  std::memcpy(hdr, head + 16, sizeof(hdr));
  if (hdr[0] != *nx || hdr[1] != *ny || hdr[2] != *nz || hdr[4] != kRestartFields ||
      hdr[5] != 4 || hdr[6] != 8) {
    std::fclose(f);
    return fail(VK_ERR_ARGUMENT, "read_restart: the record is for a " + std::to_string(hdr[0]) +
                                     "x" + std::to_string(hdr[1]) + "x" +
                                     std::to_string(hdr[2]) + " grid");
  }
  // the whole file's length before anything is written: a short file
  // leaves every caller array as it was
  const long long want = (long long)(8 + sizeof(head)) +
                         kRestartFields * (8 + (long long)(nf * sizeof(double))) +
                         4 * (8 + (long long)(ni * sizeof(double)));
  const long long here = std::ftell(f);
  if (here < 0 || std::fseek(f, 0, SEEK_END) != 0 || std::ftell(f) != want ||
      std::fseek(f, here, SEEK_SET) != 0) {
    std::fclose(f);
    return fail(VK_ERR_ARGUMENT, std::string("read_restart: short or damaged ") + path);
  }
  int rc = VK_OK;
  bool ok = true;
  for (int q = 0; q < kRestartFields && ok && !rc; q++) {
    HostOut h;
    rc = host_out(fields[q], nf, h, "read_restart: field");
    if (!rc) ok = get_record(f, h.p, nf * sizeof(double));
    if (!rc && ok) rc = h.finish("read_restart");
  }
  if (ok && !rc) {
    HostOut h;
    rc = host_out(cellvalue, 4 * ni, h, "read_restart: cellvalue");
    for (int s = 0; s < 4 && ok && !rc; s++) ok = get_record(f, h.p + s * ni, ni * sizeof(double));
    if (!rc && ok) rc = h.finish("read_restart");
  }
  std::fclose(f);
  if (rc) return rc;
  if (!ok) return fail(VK_ERR_ARGUMENT, std::string("read_restart: short or damaged ") + path);
  *l = hdr[3];
  return VK_OK;
}

}  // extern "C"
