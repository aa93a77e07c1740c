// Plumbing of the C ABI (errors, memory, streams) and the host-side Givens plan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "exc_common.hpp"

namespace exc_detail {

static thread_local std::string g_last_error;

void set_last_error(const char* what, hipError_t e, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_last_error = buf;
}

void set_last_message(const char* what, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s (%s:%d)", what, file, line);
  g_last_error = buf;
}

}  // namespace exc_detail

using exc_detail::stream;

namespace exc_prof {

namespace {
struct Rec {
  std::string name, label;
  hipEvent_t a, b;
  double flops, bytes;
};
thread_local std::string t_label;
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;
}  // namespace

bool enabled() { return g_on; }

hipEvent_t mark(hipStream_t s) {
  if (!g_on) return nullptr;
  hipEvent_t e = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_pool.empty()) {
      e = g_pool.back();
      g_pool.pop_back();
    }
  }
  if (!e && hipEventCreate(&e) != hipSuccess) return nullptr;
  hipEventRecord(e, s);
  return e;
}

void book(const char* name, hipEvent_t a, hipEvent_t b, double flops, double bytes) {
  std::string label;
  label.swap(t_label);
  if (!a || !b) return;
  std::lock_guard<std::mutex> g(g_mu);
  g_recs.push_back(Rec{name, label, a, b, flops, bytes});
}

void label_next(const char* label) { t_label = label ? label : ""; }

}  // namespace exc_prof

extern "C" {

const char* exc_last_error(void) { return exc_detail::g_last_error.c_str(); }

int exc_version(void) { return 1; }

int exc_device_count(int* n) {
  EXC_REQUIRE(n, EXC_ERR_ARG, "exc_device_count: null");
  EXC_CHECK_HIP(hipGetDeviceCount(n));
  return EXC_OK;
}

int exc_set_device(int dev) {
  EXC_CHECK_HIP(hipSetDevice(dev));
  return EXC_OK;
}

int exc_malloc(void** d_ptr, size_t bytes) {
  EXC_REQUIRE(d_ptr, EXC_ERR_ARG, "exc_malloc: null");
  *d_ptr = nullptr;
  if (bytes == 0) bytes = 256;
  EXC_CHECK_HIP(hipMalloc(d_ptr, bytes));
  return EXC_OK;
}

int exc_free(void* d_ptr) {
  if (d_ptr) EXC_CHECK_HIP(hipFree(d_ptr));
  return EXC_OK;
}

int exc_host_alloc(void** h_ptr, size_t bytes) {
  EXC_REQUIRE(h_ptr, EXC_ERR_ARG, "exc_host_alloc: null");
  if (bytes == 0) bytes = 256;
  EXC_CHECK_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault));
  return EXC_OK;
}

int exc_host_free(void* h_ptr) {
  if (h_ptr) EXC_CHECK_HIP(hipHostFree(h_ptr));
  return EXC_OK;
}

int exc_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, exc_stream_t s) {
  if (bytes == 0) return EXC_OK;
  EXC_CHECK_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, stream(s)));
  EXC_CHECK_HIP(hipStreamSynchronize(stream(s)));
  return EXC_OK;
}

int exc_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, exc_stream_t s) {
  if (bytes == 0) return EXC_OK;
  EXC_CHECK_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, stream(s)));
  EXC_CHECK_HIP(hipStreamSynchronize(stream(s)));
  return EXC_OK;
}

int exc_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, exc_stream_t s) {
  if (bytes == 0) return EXC_OK;
  EXC_CHECK_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, stream(s)));
  return EXC_OK;
}

int exc_memset(void* d_dst, int value, size_t bytes, exc_stream_t s) {
  if (bytes == 0) return EXC_OK;
  EXC_CHECK_HIP(hipMemsetAsync(d_dst, value, bytes, stream(s)));
  return EXC_OK;
}

int exc_stream_create(exc_stream_t* s) {
  EXC_REQUIRE(s, EXC_ERR_ARG, "exc_stream_create: null");
  hipStream_t h;
  EXC_CHECK_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = reinterpret_cast<exc_stream_t>(h);
  return EXC_OK;
}

int exc_stream_destroy(exc_stream_t s) {
  if (s) EXC_CHECK_HIP(hipStreamDestroy(stream(s)));
  return EXC_OK;
}

int exc_stream_sync(exc_stream_t s) {
  EXC_CHECK_HIP(hipStreamSynchronize(stream(s)));
  return EXC_OK;
}

int exc_prof_enable(int on) {
  exc_prof::g_on = on != 0;
  return EXC_OK;
}

// {"name": {"launches": n, "ms": total, "flops": total, "bytes": total}, ...}; clears the records.
int exc_prof_report(char* buf, size_t cap) {
  EXC_REQUIRE(buf && cap > 2, EXC_ERR_ARG, "exc_prof_report: bad buffer");
  std::lock_guard<std::mutex> g(exc_prof::g_mu);
  struct Agg {
    long n = 0;
    double ms = 0, flops = 0, bytes = 0;
  };
  std::vector<std::pair<std::string, Agg>> agg;
  for (auto& r : exc_prof::g_recs) {
    hipEventSynchronize(r.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, r.a, r.b);
    for (const std::string* nm : {&r.name, &r.label}) {
      if (nm->empty()) continue;
      auto it = std::find_if(agg.begin(), agg.end(), [&](auto& p) { return p.first == *nm; });
      if (it == agg.end()) {
        agg.push_back({*nm, Agg()});
        it = agg.end() - 1;
      }
      it->second.n++;
      it->second.ms += ms;
      it->second.flops += r.flops;
      it->second.bytes += r.bytes;
    }
    exc_prof::g_pool.push_back(r.a);
    exc_prof::g_pool.push_back(r.b);
  }
  exc_prof::g_recs.clear();
  std::string out = "{";
  for (size_t i = 0; i < agg.size(); i++) {
    char tmp[512];
    snprintf(tmp, sizeof tmp, "%s\"%s\": {\"launches\": %ld, \"ms\": %.6f, \"flops\": %.6e, \"bytes\": %.6e}",
             i ? ", " : "", agg[i].first.c_str(), agg[i].second.n, agg[i].second.ms, agg[i].second.flops,
             agg[i].second.bytes);
    out += tmp;
  }
  out += "}";
  EXC_REQUIRE(out.size() + 1 <= cap, EXC_ERR_SPACE, "exc_prof_report: buffer too small");
  memcpy(buf, out.c_str(), out.size() + 1);
  return EXC_OK;
}

long exc_givens_half(long K) { return (K % 2 == 0 ? K : K + 1) / 2; }

// The draw order of bbtsvd.cpp:729-779: per pass a std::shuffle of 0..K-1 over one
// default_random_engine(seed) shared by all passes, for odd K an extra index from rand() that
// differs from the last one; the unmask plan reverses each pass's list, the pass order and the
// rotation values, and each pair is stored smaller index first. c, s are two rand() draws per
// pair in [-1, 1], normalised by pow(c^2 + s^2, 0.5) on the host's C library.
int exc_givens_plan(long K, long passes, int mask, unsigned seed, int32_t* h_pairs, double* h_cs) {
  EXC_REQUIRE(K >= 2 && passes >= 0 && h_pairs && h_cs, EXC_ERR_ARG, "exc_givens_plan: bad arguments");
  const size_t k = (size_t)K;
  srand(seed);
  auto dre = std::default_random_engine(seed);
  const size_t len = (k % 2 == 0) ? k : k + 1;
  const size_t half = len / 2;
  std::vector<std::vector<size_t>> idx((size_t)passes);
  for (size_t p = 0; p < (size_t)passes; p++) {
    std::vector<size_t>& t = idx[p];
    t.reserve(len);
    for (size_t j = 0; j < k; j++) t.push_back(j);
    std::shuffle(t.begin(), t.end(), dre);
    if (k % 2 == 1) {
      t.push_back((size_t)rand() % k);
      while (t[k] == t[k - 1]) t[k] = (size_t)rand() % k;
    }
    if (!mask) std::reverse(t.begin(), t.end());
    for (size_t j = 0; j < len; j += 2)
      if (t[j] > t[j + 1]) std::swap(t[j], t[j + 1]);
  }
  if (!mask) std::reverse(idx.begin(), idx.end());
  std::vector<std::vector<double>> cc((size_t)passes), ss((size_t)passes);
  for (size_t p = 0; p < (size_t)passes; p++) {
    std::vector<double>& ci = cc[p];
    std::vector<double>& si = ss[p];
    ci.assign(half, 0.0);
    si.assign(half, 0.0);
    for (size_t j = 0; j < half; j++) {
      ci[j] = (rand() / (double)RAND_MAX) * (1.0 - -1.0) + -1.0;
      si[j] = (rand() / (double)RAND_MAX) * (1.0 - -1.0) + -1.0;
      const double root = std::pow(ci[j] * ci[j] + si[j] * si[j], 0.5);
      ci[j] /= root;
      si[j] /= root;
    }
    if (!mask) {
      std::reverse(ci.begin(), ci.end());
      std::reverse(si.begin(), si.end());
    }
  }
  if (!mask) {
    std::reverse(cc.begin(), cc.end());
    std::reverse(ss.begin(), ss.end());
  }
  for (size_t p = 0; p < (size_t)passes; p++)
    for (size_t j = 0; j < half; j++) {
      h_pairs[(p * half + j) * 2] = (int32_t)idx[p][2 * j];
      h_pairs[(p * half + j) * 2 + 1] = (int32_t)idx[p][2 * j + 1];
      h_cs[(p * half + j) * 2] = cc[p][j];
      h_cs[(p * half + j) * 2 + 1] = ss[p][j];
    }
  return EXC_OK;
}

}  // extern "C"
