// vk_runtime.cpp -- device selection, the library stream, the residency
// registry (host array -> device mirror), workspace and error reporting.
#include "vk_runtime.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <thread>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include <sys/stat.h>
#include <sys/types.h>

namespace vk {

namespace {

struct Mirror {
  size_t bytes;
  void* dev;
};

std::map<uintptr_t, Mirror>& registry() {
  static std::map<uintptr_t, Mirror> m;
  return m;
}

std::string& last_error() {
  static std::string s;
  return s;
}

int g_last_status = VK_OK;

struct Work {
  void* ptr = nullptr;
  size_t bytes = 0;
};
Work g_work[kWorkSlots];

// Guard zones: with them on, dev_malloc places a zone of kGuardBytes
// filled with kGuardByte before and after every allocation and keeps the
// allocation here, under the address it returned.
constexpr size_t kGuardBytes = 64 * 1024;
constexpr unsigned char kGuardByte = 0xA5;
struct Guarded {
  unsigned char* base;  // start of the front zone
  size_t bytes;         // the caller's bytes
  std::string what;
};
bool g_guard = false;
std::map<uintptr_t, Guarded>& guarded() {
  static std::map<uintptr_t, Guarded> m;
  return m;
}

}  // namespace

hipError_t dev_malloc(void** p, size_t bytes, const char* what) {
  if (!g_guard) return hipMalloc(p, bytes);
  void* raw = nullptr;
  hipError_t e = hipMalloc(&raw, bytes + 2 * kGuardBytes);
  if (e != hipSuccess) return e;
  unsigned char* b = static_cast<unsigned char*>(raw);
  e = hipMemset(b, kGuardByte, kGuardBytes);
  if (e == hipSuccess) e = hipMemset(b + kGuardBytes + bytes, kGuardByte, kGuardBytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(raw);
    return e;
  }
  *p = b + kGuardBytes;
  guarded()[reinterpret_cast<uintptr_t>(*p)] = Guarded{b, bytes, what};
  return hipSuccess;
}

hipError_t dev_free(void* p) {
  auto& m = guarded();
  auto it = m.find(reinterpret_cast<uintptr_t>(p));
  if (it == m.end()) return hipFree(p);
  unsigned char* base = it->second.base;
  m.erase(it);
  return hipFree(base);
}

Runtime& rt() {
  static Runtime r;
  return r;
}

int fail(int code, const std::string& msg) {
  last_error() = msg;
  g_last_status = code;
  return code;
}

int last_status() { return g_last_status; }

unsigned long long& call_epoch() {
  static unsigned long long e = 0;
  return e;
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return VK_OK;
  return fail(VK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void* resolve(const void* p, size_t bytes, const char* what) {
  if (p == nullptr) {
    fail(VK_ERR_NOT_RESIDENT, std::string(what) + ": null array");
    return nullptr;
  }
  auto& reg = registry();
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  auto it = reg.upper_bound(a);
  if (it != reg.begin()) {
    --it;
    uintptr_t base = it->first;
    if (a >= base && a < base + it->second.bytes) {
      size_t off = a - base;
      if (off + bytes > it->second.bytes) {
        fail(VK_ERR_ARGUMENT, std::string(what) + ": registered array is " +
                                      std::to_string(it->second.bytes) + " bytes, need " +
                                      std::to_string(off + bytes));
        return nullptr;
      }
      return static_cast<char*>(it->second.dev) + off;
    }
  }
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e == hipSuccess && attr.type == hipMemoryTypeDevice) {
    // a device pointer must leave room for the whole field in its
    // allocation; where the runtime cannot say (memory mapped by the
    // virtual-memory API) the caller's size is taken on trust
    size_t ext = 0;
    if (device_extent(p, &ext) && ext < bytes) {
      fail(VK_ERR_ARGUMENT, std::string(what) + ": device allocation holds " +
                                std::to_string(ext) + " bytes from this address, need " +
                                std::to_string(bytes));
      return nullptr;
    }
    return const_cast<void*>(p);
  }
  (void)hipGetLastError();  // clear the sticky error of the query
  fail(VK_ERR_NOT_RESIDENT,
       std::string(what) + ": neither a registered host array nor device memory");
  return nullptr;
}

bool is_registered(const void* p) {
  return rt().initialised && registry().count(reinterpret_cast<uintptr_t>(p)) != 0;
}

bool is_device(const void* p) {
  if (!rt().initialised || !p) return false;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice;
}

bool device_extent(const void* p, size_t* bytes) {
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  hipError_t e = hipMemGetAddressRange(&base, &size, const_cast<void*>(p));
  if (e != hipSuccess || base == nullptr) {
    (void)hipGetLastError();
    return false;
  }
  uintptr_t a = reinterpret_cast<uintptr_t>(p), b = reinterpret_cast<uintptr_t>(base);
  if (a < b || a - b > size) return false;
  *bytes = size - (a - b);
  // a guarded allocation ends where its back zone starts
  auto& m = guarded();
  auto it = m.upper_bound(a);
  if (it != m.begin()) {
    --it;
    const uintptr_t u = it->first, end = u + it->second.bytes;
    if (reinterpret_cast<uintptr_t>(it->second.base) == b) {
      *bytes = (a < u || a > end) ? 0 : end - a;  // 0 inside a zone
    }
  }
  return true;
}

std::vector<void (*)()>& finalize_hooks() {
  static std::vector<void (*)()> v;
  return v;
}

void on_finalize(void (*hook)()) {
  for (auto h : finalize_hooks())
    if (h == hook) return;
  finalize_hooks().push_back(hook);
}

void* workspace(WorkSlot slot, size_t bytes, bool* grown) {
  Work& w = g_work[slot];
  if (grown) *grown = false;
  if (w.bytes >= bytes && w.ptr) return w.ptr;
  if (w.ptr) {
    // every stream of the library may still read the old buffer
    (void)hipDeviceSynchronize();
    (void)dev_free(w.ptr);
    w.ptr = nullptr;
    w.bytes = 0;
  }
  const std::string what = "workspace slot " + std::to_string((int)slot);
  if (dev_malloc(&w.ptr, bytes, what.c_str()) != hipSuccess) {
    w.ptr = nullptr;
    fail(VK_ERR_HIP, "workspace allocation of " + std::to_string(bytes) + " bytes failed");
    return nullptr;
  }
  w.bytes = bytes;
  if (grown) *grown = true;
  return w.ptr;
}

namespace {
struct ProfSlot {
  std::string name;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t used = 0;
};
bool g_prof_on = false;
std::vector<ProfSlot>& prof_slots() {
  static std::vector<ProfSlot> v;
  return v;
}
}  // namespace

ProfScope::ProfScope(const char* name, hipStream_t stream) : s(stream) {
  if (!g_prof_on) return;
  auto& v = prof_slots();
  int idx = -1;
  for (size_t q = 0; q < v.size(); q++)
    if (v[q].name == name) idx = (int)q;
  if (idx < 0) {
    v.push_back(ProfSlot{name, {}, 0});
    idx = (int)v.size() - 1;
  }
  ProfSlot& ps = v[idx];
  if (ps.used == ps.ev.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    ps.ev.emplace_back(a, b);
  }
  if (hipEventRecord(ps.ev[ps.used].first, s) != hipSuccess) return;
  slot = idx;
}

ProfScope::~ProfScope() {
  if (slot < 0) return;
  ProfSlot& ps = prof_slots()[slot];
  if (hipEventRecord(ps.ev[ps.used].second, s) == hipSuccess) ps.used++;
}

}  // namespace vk

using namespace vk;

extern "C" {

int vk_init(const int* device) {
  int dev = device ? *device : 0;
  Runtime& r = rt();
  if (r.initialised) {
    if (r.device == dev) return VK_OK;
    return fail(VK_ERR_ARGUMENT, "vk_init: already initialised on another device");
  }
  int count = 0;
  int rc = hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (rc) return rc;
  if (dev < 0 || dev >= count)
    return fail(VK_ERR_ARGUMENT, "vk_init: device " + std::to_string(dev) + " of " +
                                     std::to_string(count));
  rc = hip_check(hipSetDevice(dev), "hipSetDevice");
  if (rc) return rc;
  {
    const char* gv = std::getenv("VK_GUARD");
    g_guard = gv && *gv && std::strcmp(gv, "0") != 0;
  }
  rc = hip_check(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking), "hipStreamCreate");
  if (rc) return rc;
  for (int q = 0; q < 3; q++) {
    rc = hip_check(hipStreamCreateWithFlags(&r.aux[q], hipStreamNonBlocking), "hipStreamCreate");
    if (rc) return rc;
    rc = hip_check(hipEventCreateWithFlags(&r.ev_join[q], hipEventDisableTiming),
                   "hipEventCreate");
    if (rc) return rc;
  }
  rc = hip_check(hipEventCreateWithFlags(&r.ev_fork, hipEventDisableTiming), "hipEventCreate");
  if (rc) return rc;
  r.device = dev;
  r.initialised = true;
  return VK_OK;
}

int vk_finalize(void) {
  Runtime& r = rt();
  if (!r.initialised) return VK_OK;
  (void)hipDeviceSynchronize();
  for (auto h : finalize_hooks()) h();
  for (auto& kv : registry()) (void)dev_free(kv.second.dev);
  registry().clear();
  for (auto& w : g_work)
    if (w.ptr) {
      (void)dev_free(w.ptr);
      w.ptr = nullptr;
      w.bytes = 0;
    }
  for (int q = 0; q < 3; q++) {
    (void)hipStreamSynchronize(r.aux[q]);
    (void)hipStreamDestroy(r.aux[q]);
    (void)hipEventDestroy(r.ev_join[q]);
    r.aux[q] = nullptr;
    r.ev_join[q] = nullptr;
  }
  (void)hipEventDestroy(r.ev_fork);
  r.ev_fork = nullptr;
  (void)hipStreamDestroy(r.stream);
  r.stream = nullptr;
  r.initialised = false;
  return VK_OK;
}

const char* vk_last_error(void) { return last_error().c_str(); }

int vk_register(const void* host, const size_t* bytes) {
  VK_REQUIRE_INIT();
  if (!host || !bytes || *bytes == 0) return fail(VK_ERR_ARGUMENT, "vk_register: empty array");
  auto& reg = registry();
  uintptr_t a = reinterpret_cast<uintptr_t>(host);
  auto it = reg.find(a);
  if (it != reg.end()) {
    if (it->second.bytes >= *bytes) return VK_OK;
    return fail(VK_ERR_ARGUMENT, "vk_register: address already registered with a smaller size");
  }
  void* d = nullptr;
  int rc = hip_check(dev_malloc(&d, *bytes, "vk_register"), "vk_register hipMalloc");
  if (rc) return rc;
  rc = hip_check(hipMemsetAsync(d, 0, *bytes, rt().stream), "vk_register memset");
  if (rc) return rc;
  reg[a] = Mirror{*bytes, d};
  return VK_OK;
}

int vk_unregister(const void* host) {
  VK_REQUIRE_INIT();
  auto& reg = registry();
  auto it = reg.find(reinterpret_cast<uintptr_t>(host));
  if (it == reg.end()) return fail(VK_ERR_ARGUMENT, "vk_unregister: not registered");
  (void)hipStreamSynchronize(rt().stream);
  (void)dev_free(it->second.dev);
  reg.erase(it);
  return VK_OK;
}

int vk_upload(const void* host) {
  VK_REQUIRE_INIT();
  auto& reg = registry();
  auto it = reg.find(reinterpret_cast<uintptr_t>(host));
  if (it == reg.end()) return fail(VK_ERR_NOT_RESIDENT, "vk_upload: array not registered");
  int rc = hip_check(hipMemcpyAsync(it->second.dev, host, it->second.bytes,
                                    hipMemcpyHostToDevice, rt().stream),
                     "vk_upload");
  if (rc) return rc;
  return hip_check(hipStreamSynchronize(rt().stream), "vk_upload sync");
}

int vk_download(void* host) {
  VK_REQUIRE_INIT();
  auto& reg = registry();
  auto it = reg.find(reinterpret_cast<uintptr_t>(host));
  if (it == reg.end()) return fail(VK_ERR_NOT_RESIDENT, "vk_download: array not registered");
  int rc = hip_check(hipMemcpyAsync(host, it->second.dev, it->second.bytes,
                                    hipMemcpyDeviceToHost, rt().stream),
                     "vk_download");
  if (rc) return rc;
  return hip_check(hipStreamSynchronize(rt().stream), "vk_download sync");
}

int vk_synchronize(void) {
  VK_REQUIRE_INIT();
  return hip_check(hipStreamSynchronize(rt().stream), "vk_synchronize");
}

int vk_malloc(void** dptr, const size_t* bytes) {
  VK_REQUIRE_INIT();
  if (!dptr || !bytes) return fail(VK_ERR_ARGUMENT, "vk_malloc: null argument");
  int rc = hip_check(dev_malloc(dptr, *bytes ? *bytes : 1, "vk_malloc"), "vk_malloc");
  if (rc) return rc;
  return hip_check(hipMemsetAsync(*dptr, 0, *bytes ? *bytes : 1, rt().stream), "vk_malloc memset");
}

int vk_free(void* dptr) {
  VK_REQUIRE_INIT();
  (void)hipStreamSynchronize(rt().stream);
  return hip_check(dev_free(dptr), "vk_free");
}

int vk_memcpy_htod(void* dst, const void* src, const size_t* bytes) {
  VK_REQUIRE_INIT();
  int rc = hip_check(hipMemcpyAsync(dst, src, *bytes, hipMemcpyHostToDevice, rt().stream),
                     "vk_memcpy_htod");
  if (rc) return rc;
  return hip_check(hipStreamSynchronize(rt().stream), "vk_memcpy_htod sync");
}

int vk_memcpy_dtoh(void* dst, const void* src, const size_t* bytes) {
  VK_REQUIRE_INIT();
  int rc = hip_check(hipMemcpyAsync(dst, src, *bytes, hipMemcpyDeviceToHost, rt().stream),
                     "vk_memcpy_dtoh");
  if (rc) return rc;
  return hip_check(hipStreamSynchronize(rt().stream), "vk_memcpy_dtoh sync");
}

int vk_profile_begin(void) {
  VK_REQUIRE_INIT();
  int rc = hip_check(hipStreamSynchronize(rt().stream), "vk_profile_begin");
  if (rc) return rc;
  for (auto& ps : prof_slots()) ps.used = 0;
  g_prof_on = true;
  rt().serial = true;
  return VK_OK;
}

int vk_profile_end(void) {
  VK_REQUIRE_INIT();
  g_prof_on = false;
  rt().serial = false;
  return hip_check(hipStreamSynchronize(rt().stream), "vk_profile_end");
}

int vk_profile_count(int* n) {
  *n = (int)prof_slots().size();
  return VK_OK;
}

int vk_profile_entry(const int* index, char* name, const int* name_len, double* total_ms,
                     int* launches) {
  auto& v = prof_slots();
  if (*index < 0 || *index >= (int)v.size()) return fail(VK_ERR_ARGUMENT, "profile index");
  ProfSlot& ps = v[*index];
  double tot = 0.0;
  for (size_t q = 0; q < ps.used; q++) {
    float ms = 0.0f;
    int rc = hip_check(hipEventElapsedTime(&ms, ps.ev[q].first, ps.ev[q].second),
                       "hipEventElapsedTime");
    if (rc) return rc;
    tot += ms;
  }
  std::snprintf(name, (size_t)*name_len, "%s", ps.name.c_str());
  *total_ms = tot;
  *launches = (int)ps.used;
  return VK_OK;
}

int vk_extent(const void* p, size_t* bytes) {
  VK_REQUIRE_INIT();
  if (!p || !bytes) return fail(VK_ERR_ARGUMENT, "vk_extent: null argument");
  auto& reg = registry();
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  auto it = reg.upper_bound(a);
  if (it != reg.begin()) {
    --it;
    if (a >= it->first && a < it->first + it->second.bytes) {
      *bytes = it->second.bytes - (a - it->first);
      return VK_OK;
    }
  }
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess || attr.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    return fail(VK_ERR_NOT_RESIDENT, "vk_extent: neither a registered host array nor device memory");
  }
  if (!device_extent(p, bytes))
    return fail(VK_ERR_ARGUMENT, "vk_extent: the runtime reports no allocation range here");
  return VK_OK;
}

int vk_guard_check(int* damaged) {
  VK_REQUIRE_INIT();
  if (damaged) *damaged = g_guard ? 0 : -1;
  if (!g_guard) return VK_OK;
  int rc = hip_check(hipDeviceSynchronize(), "vk_guard_check");
  if (rc) return rc;
  std::vector<unsigned char> h(kGuardBytes);
  std::string msg;
  int n = 0;
  for (auto& kv : guarded()) {
    const Guarded& a = kv.second;
    for (int side = 0; side < 2; side++) {
      unsigned char* z = a.base + (side ? kGuardBytes + a.bytes : 0);
      rc = hip_check(hipMemcpy(h.data(), z, kGuardBytes, hipMemcpyDeviceToHost), "vk_guard_check");
      if (rc) return rc;
      size_t first = kGuardBytes, last = 0, cnt = 0;
      for (size_t q = 0; q < kGuardBytes; q++)
        if (h[q] != kGuardByte) {
          first = q < first ? q : first;
          last = q;
          cnt++;
        }
      if (!cnt) continue;
      n++;
      // offsets from the allocation's first byte (negative: before it)
      const long long o0 = side ? (long long)(a.bytes + first) : (long long)first - (long long)kGuardBytes;
      const long long o1 = side ? (long long)(a.bytes + last) : (long long)last - (long long)kGuardBytes;
      msg += (msg.empty() ? "" : "; ") + a.what + " (" + std::to_string(a.bytes) + " bytes): " +
             std::to_string(cnt) + " bytes written " + (side ? "after its end" : "before its start") +
             ", offsets " + std::to_string(o0) + ".." + std::to_string(o1);
      rc = hip_check(hipMemset(z, kGuardByte, kGuardBytes), "vk_guard_check");
      if (rc) return rc;
    }
  }
  rc = hip_check(hipDeviceSynchronize(), "vk_guard_check");
  if (rc) return rc;
  if (damaged) *damaged = n;
  if (n) return fail(VK_ERR_GUARD, "guard zones written: " + msg);
  return VK_OK;
}

int vk_memset(void* dst, const int* value, const size_t* bytes) {
  VK_REQUIRE_INIT();
  return hip_check(hipMemsetAsync(dst, *value, *bytes, rt().stream), "vk_memset");
}


// readgrid3d (set_levelset.f:202-214): nx*ny*nz list-directed records
// "a b c phi0" of structure.dat in k, j, i order into the interior of
// phi0 (the halo is left as it is).  The list-directed rules used: values
// are separated by blanks or a comma, an item may continue on the next
// records, the rest of the record after an item's fourth value is
// skipped, and the exponent letter may be E or D.  Repeat counts (r*c),
// null values and the slash terminator are refused.  phi0 is host memory,
// as in the reference; when it is registered its device mirror is
// refreshed.

namespace {

enum ItemStatus { kItemOk = 0, kItemEnd, kItemNull, kItemRepeat, kItemBad };

// One item from p: its four values into val, p left at the start of the
// next item's record.  `multi` reports an item that took more than one
// record (or skipped a blank one).
ItemStatus parse_item(const char*& p, const char* end, double val[4], bool* multi) {
  char tok[128];
  int nv = 0;
  // a comma before the first value of an item is a null value, as is a
  // comma after another comma (blanks and record ends between them do not
  // count)
  bool sep_comma = true;
  *multi = false;
  while (nv < 4) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == ',')) {
      if (*p == ',') {
        if (sep_comma) return kItemNull;
        sep_comma = true;
      }
      p++;
    }
    if (p >= end) return kItemEnd;
    if (*p == '\n') {  // the item continues on the next record
      *multi = true;
      p++;
      continue;
    }
    const char* t0 = p;
    while (p < end && *p != ' ' && *p != '\t' && *p != '\r' && *p != ',' && *p != '\n') p++;
    const size_t n = (size_t)(p - t0);
    if (n == 0 || n >= sizeof(tok)) return kItemBad;
    for (size_t q = 0; q < n; q++) {
      const char c = t0[q];
      if (c == '*' || c == '/') return kItemRepeat;
      tok[q] = (c == 'd' || c == 'D') ? 'e' : c;
    }
    tok[n] = '\0';
    char* e = nullptr;
    const double v = std::strtod(tok, &e);
    if (e != tok + n) return kItemBad;
    val[nv++] = v;
    sep_comma = false;
  }
  while (p < end && *p != '\n') p++;  // the rest of the record
  if (p < end) p++;
  return kItemOk;
}

int item_fail(ItemStatus st) {
  switch (st) {
    case kItemEnd: return fail(VK_ERR_ARGUMENT, "readgrid3d: structure.dat ends early");
    case kItemNull: return fail(VK_ERR_ARGUMENT, "readgrid3d: null value");
    case kItemRepeat:
      return fail(VK_ERR_ARGUMENT, "readgrid3d: repeat counts and '/' not supported");
    default: return fail(VK_ERR_ARGUMENT, "readgrid3d: bad value in structure.dat");
  }
}

// The file as one item per record, parsed in parallel: the text split on
// record boundaries, each chunk's items parsed by parse_item.  Returns
// false (the caller then reads serially) for anything else -- an item over
// several records, a blank record, any refusal -- so that the serial
// reader decides those cases.
bool readgrid3d_parallel(const std::string& text, long long nitems, std::vector<double>& out) {
  const char* const base = text.data();
  const char* const end = base + text.size();
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<long long>(std::min<unsigned>(hw, 16),
                                          1 + (long long)text.size() / (1 << 20));
  if (nt < 2) return false;
  std::vector<const char*> cut((size_t)nt + 1);
  cut[0] = base;
  cut[nt] = end;
  for (int t = 1; t < nt; t++) {
    const char* c = base + text.size() * (size_t)t / (size_t)nt;
    while (c < end && *c != '\n') c++;
    cut[t] = c < end ? c + 1 : end;
  }
  std::vector<std::vector<double>> vals((size_t)nt);
  std::vector<char> ok((size_t)nt, 1);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      const char* p = cut[t];
      const char* e = cut[t + 1];
      double v[4];
      while (p < e) {
        bool multi = false;
        if (parse_item(p, e, v, &multi) != kItemOk || multi) {
          ok[t] = 0;
          return;
        }
        vals[t].push_back(v[3]);
      }
    });
  for (auto& x : th) x.join();
  long long total = 0;
  for (int t = 0; t < nt; t++) {
    if (!ok[t] && total < nitems) return false;  // irregular before the last item
    total += (long long)vals[t].size();
    if (!ok[t]) break;
  }
  if (total < nitems) return false;
  out.resize((size_t)nitems);
  long long q = 0;
  for (int t = 0; t < nt && q < nitems; t++)
    for (size_t r = 0; r < vals[t].size() && q < nitems; r++) out[(size_t)q++] = vals[t][r];
  return true;
}

}  // namespace

int vk_readgrid3d(const int* nx, const int* ny, const int* nz, double* phi0) {
  if (*nx < 1 || *ny < 1 || *nz < 1) return fail(VK_ERR_BAD_SIZE, "readgrid3d: nx, ny, nz >= 1");
  if (!phi0) return fail(VK_ERR_ARGUMENT, "readgrid3d: phi0 is null");
  std::FILE* f = std::fopen("structure.dat", "rb");
  if (!f) return fail(VK_ERR_ARGUMENT, "readgrid3d: cannot open structure.dat");
  std::string text;
  {
    // one read of the whole file where its size is known, else in pieces
    struct stat sb;
    if (fstat(fileno(f), &sb) == 0 && sb.st_size > 0) {
      text.resize((size_t)sb.st_size);
      text.resize(std::fread(&text[0], 1, text.size(), f));
    }
    std::vector<char> buf(1 << 20);
    size_t got;
    while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) text.append(buf.data(), got);
    std::fclose(f);
  }
  const size_t sy = (size_t)*nx + 2, sz = sy * ((size_t)*ny + 2);
  const long long nitems = (long long)*nx * *ny * *nz;
  std::vector<double> fast;
  if (readgrid3d_parallel(text, nitems, fast)) {
    long long q = 0;
    for (int k = 1; k <= *nz; k++)
      for (int j = 1; j <= *ny; j++)
        for (int i = 1; i <= *nx; i++) phi0[(size_t)i + sy * (size_t)j + sz * (size_t)k] = fast[(size_t)q++];
  } else {
    const char* p = text.data();
    const char* const end = p + text.size();
    for (int k = 1; k <= *nz; k++)
      for (int j = 1; j <= *ny; j++)
        for (int i = 1; i <= *nx; i++) {
          double val[4];
          bool multi = false;
          const ItemStatus st = parse_item(p, end, val, &multi);
          if (st != kItemOk) return item_fail(st);
          phi0[(size_t)i + sy * (size_t)j + sz * (size_t)k] = val[3];
        }
  }
  if (rt().initialised &&
      registry().count(reinterpret_cast<uintptr_t>(static_cast<const void*>(phi0))))
    return vk_upload(phi0);
  return VK_OK;
}


// ------------------------------------------------------------ field files
// Fortran sequential unformatted records: a 4-byte length, the data, the
// length again (the layout flang and gfortran write for records under
// 2 GiB).

// a registered host array's device copy brought back before it is written
// to a file, or refreshed after it is read from one
static int host_sync(const void* host, bool down) {
  if (!rt().initialised || !registry().count(reinterpret_cast<uintptr_t>(host))) return VK_OK;
  return down ? vk_download(const_cast<void*>(host)) : vk_upload(host);
}

struct RecordWriter {
  std::FILE* f;
  std::vector<char> buf;
  bool ok = true;
  explicit RecordWriter(std::FILE* f_) : f(f_) { buf.reserve(1 << 22); }
  void record(const double* v, int n) {
    const int32_t len = (int32_t)(n * sizeof(double));
    const char* l = reinterpret_cast<const char*>(&len);
    buf.insert(buf.end(), l, l + 4);
    const char* d = reinterpret_cast<const char*>(v);
    buf.insert(buf.end(), d, d + len);
    buf.insert(buf.end(), l, l + 4);
    if (buf.size() >= (1u << 22)) flush();
  }
  void flush() {
    if (!buf.empty() && std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) ok = false;
    buf.clear();
  }
};

// OUTFIELD (nse_dominik.f:2726-2766): field.u holds the record tid, then
// one record (u, v, w, p, T, phi, u_ta, v_ta, w_ta) per halo point in k, j,
// i order; ./nseoutput2/grid.u holds x(1..nx), y(1..ny), z(1..nz), one
// record each.
int vk_outfield(const double* x, const double* y, const double* z, const double* u,
                const double* v, const double* w, const double* p, const int* nx,
                const int* ny, const int* nz, const double* tid, const double* t,
                const double* phi, const double* u_ta, const double* v_ta, const double* w_ta) {
  if (*nx < 1 || *ny < 1 || *nz < 1) return fail(VK_ERR_BAD_SIZE, "outfield: nx, ny, nz >= 1");
  const double* f9[9] = {u, v, w, p, t, phi, u_ta, v_ta, w_ta};
  for (const double* a : f9) {
    if (!a) return fail(VK_ERR_ARGUMENT, "outfield: a field is null");
    if (int rc = host_sync(a, true)) return rc;
  }
  const size_t n = ((size_t)*nx + 2) * ((size_t)*ny + 2) * ((size_t)*nz + 2);
  std::FILE* f = std::fopen("field.u", "wb");
  if (!f) return fail(VK_ERR_ARGUMENT, "outfield: cannot open field.u");
  RecordWriter wr(f);
  wr.record(tid, 1);
  double rec[9];
  for (size_t q = 0; q < n; q++) {
    for (int c = 0; c < 9; c++) rec[c] = f9[c][q];
    wr.record(rec, 9);
  }
  wr.flush();
  const bool ok = wr.ok && std::fclose(f) == 0;
  if (!ok) return fail(VK_ERR_ARGUMENT, "outfield: writing field.u failed");
  std::FILE* gfile = std::fopen("./nseoutput2/grid.u", "wb");
  if (!gfile) return fail(VK_ERR_ARGUMENT, "outfield: cannot open ./nseoutput2/grid.u");
  RecordWriter gw(gfile);
  for (int i = 1; i <= *nx; i++) gw.record(x + i, 1);
  for (int j = 1; j <= *ny; j++) gw.record(y + j, 1);
  for (int k = 1; k <= *nz; k++) gw.record(z + k, 1);
  gw.flush();
  if (!gw.ok || std::fclose(gfile) != 0) return fail(VK_ERR_ARGUMENT, "outfield: writing grid.u failed");
  return VK_OK;
}

// READFIELD (nse_dominik.f:2769-2793): ./nseoutput2/field.u, one record of
// 9 values per halo point in k, j, i order, as OUTPUT writes it.  A leading
// 8-byte record (the tid that OUTFIELD writes first, which the reference's
// read does not expect) is skipped, so OUTFIELD's field.u reads back too.
int vk_readfield(double* u, double* v, double* w, double* p, const int* nx, const int* ny,
                 const int* nz, double* t, double* phi, double* u_ta, double* v_ta,
                 double* w_ta) {
  if (*nx < 1 || *ny < 1 || *nz < 1) return fail(VK_ERR_BAD_SIZE, "readfield: nx, ny, nz >= 1");
  double* f9[9] = {u, v, w, p, t, phi, u_ta, v_ta, w_ta};
  for (double* a : f9)
    if (!a) return fail(VK_ERR_ARGUMENT, "readfield: a field is null");
  std::FILE* f = std::fopen("./nseoutput2/field.u", "rb");
  if (!f) return fail(VK_ERR_ARGUMENT, "readfield: cannot open ./nseoutput2/field.u");
  std::string text;
  {
    std::vector<char> buf(1 << 22);
    size_t got;
    while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) text.append(buf.data(), got);
    std::fclose(f);
  }
  const size_t n = ((size_t)*nx + 2) * ((size_t)*ny + 2) * ((size_t)*nz + 2);
  const char* c = text.data();
  size_t left = text.size();
  auto len_at = [&](const char* q) {
    int32_t l;
    std::memcpy(&l, q, 4);
    return l;
  };
  if (left >= 16 && len_at(c) == 8 && len_at(c + 12) == 8) {  // the tid record
    c += 16;
    left -= 16;
  }
  const size_t recb = 4 + 72 + 4;
  if (left < n * recb) return fail(VK_ERR_ARGUMENT, "readfield: field.u is too short");
  for (size_t q = 0; q < n; q++, c += recb) {
    if (len_at(c) != 72 || len_at(c + 76) != 72)
      return fail(VK_ERR_ARGUMENT, "readfield: field.u record is not 9 doubles");
    double rec[9];
    std::memcpy(rec, c + 4, 72);
    for (int k = 0; k < 9; k++) f9[k][q] = rec[k];
  }
  for (double* a : f9)
    if (int rc = host_sync(a, false)) return rc;
  return VK_OK;
}


// OUTPUT (nse_dominik.f:2606-2723): creates nseoutput2/ when missing, then
// writes, one record per interior point in k, j, i order, velocity.dat
// (the face averages ua, va, wa, each set to 0 below 1e-30 in magnitude),
// pressure, temperature, divergence, phi and cellvalue1..4; gridx/y/z.dat
// and gridnumber.dat (nx, ny, nz as 4-byte integers); field.u and grid.u as
// OUTFIELD writes them but without the tid record (the file READFIELD
// reads); and temperature_check.dat, y(j) and T(nx, j, nz/2) for j =
// 0..ny+1 in list-directed text (the values to 17 significant digits).
int vk_output(const double* u, const double* v, const double* w, const double* p,
              const double* x, const double* y, const double* z, const int* nx, const int* ny,
              const int* nz, const double* t, const double* cellvalue, const double* u_ta,
              const double* v_ta, const double* w_ta, const double* phi, const double* divu) {
  if (*nx < 1 || *ny < 1 || *nz < 1) return fail(VK_ERR_BAD_SIZE, "output: nx, ny, nz >= 1");
  const double* fl[11] = {u, v, w, p, t, cellvalue, u_ta, v_ta, w_ta, phi, divu};
  for (const double* a : fl) {
    if (!a) return fail(VK_ERR_ARGUMENT, "output: an array is null");
    if (int rc = host_sync(a, true)) return rc;
  }
  struct stat sb;
  if (stat("nseoutput2", &sb) != 0) {
    std::printf("Create nseoutput2 directory\n");
    std::fflush(stdout);
    if (mkdir("nseoutput2", 0777) != 0)
      return fail(VK_ERR_ARGUMENT, "output: cannot create nseoutput2");
  }
  static const char* names[12] = {"velocity", "temperature", "pressure", "divergence",
                                  "phi", "cellvalue1", "cellvalue2", "cellvalue3",
                                  "cellvalue4", "gridx", "gridy", "gridz"};
  std::FILE* f[12];
  for (int q = 0; q < 12; q++) {
    const std::string path = std::string("./nseoutput2/") + names[q] + ".dat";
    f[q] = std::fopen(path.c_str(), "wb");
    if (!f[q]) {
      for (int r = 0; r < q; r++) std::fclose(f[r]);
      return fail(VK_ERR_ARGUMENT, "output: cannot open a file in nseoutput2");
    }
  }
  std::vector<RecordWriter> wr;
  wr.reserve(12);
  for (int q = 0; q < 12; q++) wr.emplace_back(f[q]);
  const size_t sy = (size_t)*nx + 2, sz = sy * ((size_t)*ny + 2);
  const size_t ni = (size_t)*nx * *ny * *nz;
  auto at = [&](int i, int j, int k) { return (size_t)i + sy * (size_t)j + sz * (size_t)k; };
  for (int k = 1; k <= *nz; k++)
    for (int j = 1; j <= *ny; j++)
      for (int i = 1; i <= *nx; i++) {
        const size_t o = at(i, j, k);
        double ua = (u[at(i - 1, j, k)] + u[o]) * 0.50;
        double va = (v[at(i, j - 1, k)] + v[o]) * 0.50;
        double wa = (w[at(i, j, k - 1)] + w[o]) * 0.50;
        if (std::fabs(ua) < 1.0e-30) ua = 0.0;
        if (std::fabs(va) < 1.0e-30) va = 0.0;
        if (std::fabs(wa) < 1.0e-30) wa = 0.0;
        const double vel[3] = {ua, va, wa};
        wr[0].record(vel, 3);
        wr[1].record(t + o, 1);
        wr[2].record(p + o, 1);
        wr[3].record(divu + o, 1);
        wr[4].record(phi + o, 1);
        const size_t c = (size_t)(i - 1) + (size_t)*nx * ((size_t)(j - 1) + (size_t)*ny * (k - 1));
        for (int n = 0; n < 4; n++) wr[5 + n].record(cellvalue + c + ni * n, 1);
      }
  for (int i = 1; i <= *nx; i++) wr[9].record(x + i, 1);
  for (int j = 1; j <= *ny; j++) wr[10].record(y + j, 1);
  for (int k = 1; k <= *nz; k++) wr[11].record(z + k, 1);
  bool ok = true;
  for (int q = 0; q < 12; q++) {
    wr[q].flush();
    ok = ok && wr[q].ok;
    ok = (std::fclose(f[q]) == 0) && ok;
  }
  {
    std::FILE* gn = std::fopen("./nseoutput2/gridnumber.dat", "wb");
    if (!gn) return fail(VK_ERR_ARGUMENT, "output: cannot open gridnumber.dat");
    const int32_t len = 12, dims[3] = {*nx, *ny, *nz};
    ok = std::fwrite(&len, 4, 1, gn) == 1 && ok;
    ok = std::fwrite(dims, 4, 3, gn) == 3 && ok;
    ok = std::fwrite(&len, 4, 1, gn) == 1 && ok;
    ok = (std::fclose(gn) == 0) && ok;
  }
  {
    std::FILE* tc = std::fopen("temperature_check.dat", "w");
    if (!tc) return fail(VK_ERR_ARGUMENT, "output: cannot open temperature_check.dat");
    const int i = *nx, k = *nz / 2;
    for (int j = 0; j <= *ny + 1; j++)
      ok = std::fprintf(tc, " %.17g %.17g\n", y[j], t[at(i, j, k)]) > 0 && ok;
    ok = (std::fclose(tc) == 0) && ok;
  }
  if (!ok) return fail(VK_ERR_ARGUMENT, "output: writing failed");
  // the restart pair: field.u without the tid record, and grid.u
  std::FILE* ff = std::fopen("./nseoutput2/field.u", "wb");
  if (!ff) return fail(VK_ERR_ARGUMENT, "output: cannot open nseoutput2/field.u");
  RecordWriter fw(ff);
  const double* f9[9] = {u, v, w, p, t, phi, u_ta, v_ta, w_ta};
  const size_t n = sz * ((size_t)*nz + 2);
  double rec[9];
  for (size_t q = 0; q < n; q++) {
    for (int c = 0; c < 9; c++) rec[c] = f9[c][q];
    fw.record(rec, 9);
  }
  fw.flush();
  if (!fw.ok || std::fclose(ff) != 0) return fail(VK_ERR_ARGUMENT, "output: writing field.u failed");
  std::FILE* gf = std::fopen("./nseoutput2/grid.u", "wb");
  if (!gf) return fail(VK_ERR_ARGUMENT, "output: cannot open nseoutput2/grid.u");
  RecordWriter gw(gf);
  for (int i = 1; i <= *nx; i++) gw.record(x + i, 1);
  for (int j = 1; j <= *ny; j++) gw.record(y + j, 1);
  for (int k = 1; k <= *nz; k++) gw.record(z + k, 1);
  gw.flush();
  if (!gw.ok || std::fclose(gf) != 0) return fail(VK_ERR_ARGUMENT, "output: writing grid.u failed");
  return VK_OK;
}

}  // extern "C"
