// Host graph ingest, file formats and seeded synthetic workloads.
//
// C ABI declared in include/mi355graph.h.  Built with g++ into
// libmi355graph.so; it has no GPU code and no dependency on the HIP runtime.
#include "mi355graph.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

struct mi355g_graph {
  int64_t rows = 0, cols = 0;
  int32_t weight_type = MI355G_W_F32;
  int32_t orientation = MI355G_OUT;
  uint64_t declared_edges = 0, loaded_edges = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<uint32_t> w;
  int64_t nnz() const { return (int64_t)col.size(); }
};

// RIPSIFM1 sidecar held in memory: the header counts, the string table (one
// byte buffer plus offsets) and every table as flat u64 records.
struct mi355g_ifm {
  mi355g_ifm_header_t h{};
  std::vector<char> str_bytes;
  std::vector<uint64_t> str_off;  // string_count + 1
  std::vector<uint64_t> node_ids, edge_attrs, pips, site_pins;
  std::vector<uint64_t> nets, src_off, sources, sink_off, sinks;
  std::vector<uint64_t> cells, lnets, ports;
  std::vector<uint64_t> blocked, sink_stop;
  std::vector<uint8_t> physical, logical;
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(MI355G_ENOMEM, "out of host memory");
  } catch (const std::exception& e) {
    return fail(MI355G_EINVAL, e.what());
  }
}

// ---- counter-based RNG -------------------------------------------------
// SplitMix64 finaliser (Steele, Lea, Flood 2014) applied to a key derived
// from (seed, stream) plus the golden-ratio increment times the index.
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t index) {
  const uint64_t key = mix64(seed + 0x9E3779B97F4A7C15ull * (stream + 1));
  return mix64(key + 0x9E3779B97F4A7C15ull * (index + 1));
}
inline float unit_f32(uint64_t x) {
  // (u24 + 1) * 2^-24 lies in (0, 1] and is exact in fp32.
  return (float)((x >> 40) + 1) * (1.0f / 16777216.0f);
}
inline double open_unit_f64(uint32_t x) {  // (0, 1]
  return ((double)x + 1.0) * (1.0 / 4294967296.0);
}
// Multiply-shift range reduction of a 64-bit draw (bias below 2^-32 for the
// ranges used here).
inline uint64_t range_u64(uint64_t x, uint64_t range) {
  return (uint64_t)(((unsigned __int128)x * range) >> 64);
}

unsigned hw_threads() {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 4;
  return std::min(t, 16u);
}

template <typename F>
void parallel_for(int64_t n, F&& f) {
  const unsigned T = hw_threads();
  if (n < 1 << 16 || T == 1) {
    f(0, n, 0u);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; ++t) {
    const int64_t b = n * t / T, e = n * (t + 1) / T;
    th.emplace_back([&f, b, e, t] { f(b, e, t); });
  }
  for (auto& x : th) x.join();
}

// LSD radix sort of 64-bit keys on the low `bits` bits.
void radix_sort_u64(std::vector<uint64_t>& a, int bits) {
  std::vector<uint64_t> tmp(a.size());
  for (int shift = 0; shift < bits; shift += 11) {
    size_t cnt[2049] = {0};
    for (uint64_t x : a) cnt[((x >> shift) & 2047) + 1]++;
    for (int i = 0; i < 2048; ++i) cnt[i + 1] += cnt[i];
    for (uint64_t x : a) tmp[cnt[(x >> shift) & 2047]++] = x;
    a.swap(tmp);
  }
}

// Build a CSR from an edge list: rows by counting sort, each row ordered by
// column (stable among equal columns), optional min-reduction of duplicates.
mi355g_graph* build_from_edges(int64_t n, int64_t m, const int32_t* src,
                               const int32_t* dst, const uint32_t* w,
                               int32_t wt, int32_t orient, bool dedup) {
  auto g = new mi355g_graph();
  g->rows = g->cols = n;
  g->weight_type = wt;
  g->orientation = orient;
  g->loaded_edges = g->declared_edges = (uint64_t)m;
  const int32_t* row_of = orient == MI355G_OUT ? src : dst;
  const int32_t* col_of = orient == MI355G_OUT ? dst : src;
  std::vector<int64_t> cnt(n + 1, 0);
  for (int64_t e = 0; e < m; ++e) cnt[row_of[e] + 1]++;
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int64_t> order(m);
  {
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < m; ++e) order[cur[row_of[e]]++] = e;
  }
  g->rowptr.assign(n + 1, 0);
  g->col.reserve(m);
  g->w.reserve(m);
  for (int64_t r = 0; r < n; ++r) {
    auto b = order.begin() + cnt[r], e = order.begin() + cnt[r + 1];
    std::stable_sort(b, e, [&](int64_t x, int64_t y) { return col_of[x] < col_of[y]; });
    for (auto it = b; it != e;) {
      const int32_t c = col_of[*it];
      uint32_t best = w[*it];
      auto jt = it + 1;
      if (dedup) {
        for (; jt != e && col_of[*jt] == c; ++jt) {
          const uint32_t x = w[*jt];
          if (wt == MI355G_W_F32) {
            float fx, fb;
            std::memcpy(&fx, &x, 4);
            std::memcpy(&fb, &best, 4);
            if (fx < fb) best = x;
          } else if (x < best) {
            best = x;
          }
        }
      }
      g->col.push_back(c);
      g->w.push_back(best);
      it = jt;
    }
    g->rowptr[r + 1] = (int64_t)g->col.size();
  }
  return g;
}

mi355g_graph* transpose_graph(const mi355g_graph& a) {
  auto g = new mi355g_graph();
  g->rows = a.cols;
  g->cols = a.rows;
  g->weight_type = a.weight_type;
  g->orientation = a.orientation == MI355G_OUT ? MI355G_IN : MI355G_OUT;
  g->declared_edges = a.declared_edges;
  g->loaded_edges = a.loaded_edges;
  const int64_t m = a.nnz();
  g->rowptr.assign(g->rows + 1, 0);
  g->col.resize(m);
  g->w.resize(m);
  for (int64_t e = 0; e < m; ++e) g->rowptr[a.col[e] + 1]++;
  for (int64_t i = 0; i < g->rows; ++i) g->rowptr[i + 1] += g->rowptr[i];
  std::vector<int64_t> cur(g->rowptr.begin(), g->rowptr.end() - 1);
  for (int64_t r = 0; r < a.rows; ++r)
    for (int64_t e = a.rowptr[r]; e < a.rowptr[r + 1]; ++e) {
      const int64_t p = cur[a.col[e]]++;
      g->col[p] = (int32_t)r;
      g->w[p] = a.w[e];
    }
  return g;
}

// ---- SHA-256 (FIPS 180-4) ----------------------------------------------
struct Sha256 {
  uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                   0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  unsigned char buf[64];
  size_t blen = 0;
  uint64_t total = 0;
  static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
  void block(const unsigned char* p) {
    static const uint32_t k[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4,
        0xab1c5ed5, 0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe,
        0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f,
        0x4a7484aa, 0x5cb0a9dc, 0x76f988da, 0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7,
        0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc,
        0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b,
        0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070, 0x19a4c116,
        0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7,
        0xc67178f2};
    uint32_t wv[64];
    for (int i = 0; i < 16; ++i)
      wv[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 |
              (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
      const uint32_t s0 = rotr(wv[i - 15], 7) ^ rotr(wv[i - 15], 18) ^ (wv[i - 15] >> 3);
      const uint32_t s1 = rotr(wv[i - 2], 17) ^ rotr(wv[i - 2], 19) ^ (wv[i - 2] >> 10);
      wv[i] = wv[i - 16] + s0 + wv[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; ++i) {
      const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
      const uint32_t ch = (e & f) ^ (~e & g);
      const uint32_t t1 = hh + S1 + ch + k[i] + wv[i];
      const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
      const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
      const uint32_t t2 = S0 + mj;
      hh = g;
      g = f;
      f = e;
      e = d + t1;
      d = c;
      c = b;
      b = a;
      a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d;
    h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
  }
  void update(const void* data, size_t len) {
    const unsigned char* p = (const unsigned char*)data;
    total += len;
    while (len > 0) {
      if (blen == 0 && len >= 64) {
        block(p);
        p += 64;
        len -= 64;
        continue;
      }
      const size_t take = std::min(len, 64 - blen);
      std::memcpy(buf + blen, p, take);
      blen += take;
      p += take;
      len -= take;
      if (blen == 64) {
        block(buf);
        blen = 0;
      }
    }
  }
  void final_hex(char* out) {
    const uint64_t bits = total * 8;
    const unsigned char pad = 0x80, zero = 0;
    update(&pad, 1);
    while (blen != 56) update(&zero, 1);
    unsigned char lenb[8];
    for (int i = 0; i < 8; ++i) lenb[i] = (unsigned char)(bits >> (56 - 8 * i));
    update(lenb, 8);
    for (int i = 0; i < 8; ++i) std::snprintf(out + 8 * i, 9, "%08x", h[i]);
  }
};

// ---- binary I/O helpers --------------------------------------------------
struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) { f = std::fopen(path, mode); }
  ~File() {
    if (f) std::fclose(f);
  }
};

template <typename T>
bool read_arr(FILE* f, T* p, size_t n) {
  return n == 0 || std::fread(p, sizeof(T), n, f) == n;
}
template <typename T>
bool write_arr(FILE* f, const T* p, size_t n) {
  return n == 0 || std::fwrite(p, sizeof(T), n, f) == n;
}

// Bytes in the file (the position is restored); -1 when it cannot be sized.
int64_t file_bytes(FILE* f) {
  const long pos = std::ftell(f);
  if (pos < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
  const long size = std::ftell(f);
  if (std::fseek(f, pos, SEEK_SET) != 0) return -1;
  return size;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* mi355g_last_error(void) { return g_last_error.c_str(); }

uint64_t mi355g_rng_u64(uint64_t seed, uint64_t stream, uint64_t index) {
  return rng_u64(seed, stream, index);
}
float mi355g_rng_unit_f32(uint64_t seed, uint64_t stream, uint64_t index) {
  return unit_f32(rng_u64(seed, stream, index));
}
void mi355g_fill_unit_f32(uint64_t seed, uint64_t stream, uint64_t first, int64_t count,
                          float* buf) {
  parallel_for(count, [&](int64_t b, int64_t e, unsigned) {
    for (int64_t i = b; i < e; ++i) buf[i] = unit_f32(rng_u64(seed, stream, first + i));
  });
}

int mi355g_view(const mi355g_graph* g, mi355g_view_t* out) {
  if (!g || !out) return fail(MI355G_EINVAL, "null argument");
  out->rows = g->rows;
  out->cols = g->cols;
  out->nnz = g->nnz();
  out->weight_type = g->weight_type;
  out->orientation = g->orientation;
  out->rowptr = g->rowptr.data();
  out->col = g->col.data();
  out->w = g->w.data();
  out->declared_edges = g->declared_edges;
  out->loaded_edges = g->loaded_edges;
  return 0;
}

void mi355g_free(mi355g_graph* g) { delete g; }

int mi355g_from_csr(int64_t rows, int64_t cols, int64_t nnz, const int64_t* rowptr,
                    const int32_t* col, const uint32_t* w, int32_t weight_type,
                    int32_t orientation, mi355g_graph** out) {
  return guarded([&] {
    if (!out || rows < 0 || cols < 0 || nnz < 0 || !rowptr || (nnz > 0 && (!col || !w)))
      return fail(MI355G_EINVAL, "invalid CSR arguments");
    if (rowptr[0] != 0 || rowptr[rows] != nnz)
      return fail(MI355G_EINVAL, "rowptr must start at 0 and end at nnz");
    for (int64_t r = 0; r < rows; ++r)
      if (rowptr[r + 1] < rowptr[r]) return fail(MI355G_EINVAL, "rowptr must be monotone");
    for (int64_t e = 0; e < nnz; ++e)
      if (col[e] < 0 || col[e] >= cols) return fail(MI355G_EINVAL, "column index out of range");
    auto g = new mi355g_graph();
    g->rows = rows;
    g->cols = cols;
    g->weight_type = weight_type;
    g->orientation = orientation;
    g->declared_edges = g->loaded_edges = (uint64_t)nnz;
    g->rowptr.assign(rowptr, rowptr + rows + 1);
    g->col.assign(col, col + nnz);
    g->w.assign(w, w + nnz);
    *out = g;
    return 0;
  });
}

int mi355g_from_edges(int64_t n, int64_t m, const int32_t* src, const int32_t* dst,
                      const uint32_t* w, int32_t weight_type, int32_t orientation,
                      int32_t dedup_min, mi355g_graph** out) {
  return guarded([&] {
    if (!out || n <= 0 || m < 0 || (m > 0 && (!src || !dst || !w)))
      return fail(MI355G_EINVAL, "invalid edge-list arguments");
    for (int64_t e = 0; e < m; ++e)
      if (src[e] < 0 || src[e] >= n || dst[e] < 0 || dst[e] >= n)
        return fail(MI355G_EINVAL, "edge endpoint out of range");
    *out = build_from_edges(n, m, src, dst, w, weight_type, orientation, dedup_min != 0);
    return 0;
  });
}

int mi355g_transpose(const mi355g_graph* g, mi355g_graph** out) {
  return guarded([&] {
    if (!g || !out) return fail(MI355G_EINVAL, "null argument");
    *out = transpose_graph(*g);
    return 0;
  });
}

// ---- RIPSCSR1 ------------------------------------------------------------
int mi355g_read_ripscsr1(const char* path, mi355g_graph** out) {
  return guarded([&] {
    File f(path, "rb");
    if (!f.f) return fail(MI355G_EIO, std::string("cannot open ") + path);
    char magic[8];
    uint64_t h[10];
    if (!read_arr(f.f, magic, 8) || std::memcmp(magic, "RIPSCSR1", 8) != 0)
      return fail(MI355G_EFORMAT, "bad RIPSCSR1 magic");
    if (!read_arr(f.f, h, 10)) return fail(MI355G_EFORMAT, "truncated RIPSCSR1 header");
    // h: version, orientation, rows, cols, declared, loaded, nnz, rowptr_n, colind_n, values_n
    if (h[0] != 1) return fail(MI355G_EFORMAT, "unsupported RIPSCSR1 version");
    if (h[1] != 1) return fail(MI355G_EFORMAT, "RIPSCSR1 orientation must be 1 (incoming)");
    // vertex ids are 32-bit: this also keeps rows + 1 below wrap-around
    if (h[2] > (uint64_t)std::numeric_limits<int32_t>::max())
      return fail(MI355G_EFORMAT, "RIPSCSR1 row count exceeds 32-bit vertex ids");
    if (h[7] != h[2] + 1 || h[8] != h[6] || h[9] != h[6])
      return fail(MI355G_EFORMAT, "inconsistent RIPSCSR1 array counts");
    // the arrays must be in the file before anything is allocated for them
    const int64_t size = file_bytes(f.f);
    if (size < 0) return fail(MI355G_EIO, "cannot size the file");
    const uint64_t avail = (uint64_t)size - 88u;  // header: magic + 10 words
    if (h[7] > avail / 8 || h[6] > avail / 8 || 8 * h[7] + 8 * h[6] > avail)
      return fail(MI355G_EFORMAT, "truncated RIPSCSR1 arrays");
    if (h[3] > (uint64_t)std::numeric_limits<int32_t>::max() + 1u)
      return fail(MI355G_EFORMAT, "RIPSCSR1 column count exceeds 32-bit indices");
    auto g = new mi355g_graph();
    g->rows = (int64_t)h[2];
    g->cols = (int64_t)h[3];
    g->declared_edges = h[4];
    g->loaded_edges = h[5];
    g->weight_type = MI355G_W_F32;
    g->orientation = MI355G_IN;
    g->rowptr.resize(h[7]);
    g->col.resize(h[6]);
    g->w.resize(h[6]);
    if (!read_arr(f.f, g->rowptr.data(), h[7]) || !read_arr(f.f, g->col.data(), h[6]) ||
        !read_arr(f.f, g->w.data(), h[6])) {
      delete g;
      return fail(MI355G_EFORMAT, "truncated RIPSCSR1 arrays");
    }
    if (g->rowptr[0] != 0 || g->rowptr[g->rows] != (int64_t)h[6]) {
      delete g;
      return fail(MI355G_EFORMAT, "RIPSCSR1 rowptr must span [0, nnz]");
    }
    for (int64_t r = 0; r < g->rows; ++r)
      if (g->rowptr[r + 1] < g->rowptr[r]) {
        delete g;
        return fail(MI355G_EFORMAT, "RIPSCSR1 rowptr decreases at row " + std::to_string(r));
      }
    for (const int32_t c : g->col)
      if (c < 0 || (int64_t)c >= g->cols) {
        delete g;
        return fail(MI355G_EFORMAT, "RIPSCSR1 column index out of range");
      }
    *out = g;
    return 0;
  });
}

int mi355g_write_ripscsr1(const mi355g_graph* g, const char* path) {
  return guarded([&] {
    if (!g) return fail(MI355G_EINVAL, "null graph");
    if (g->orientation != MI355G_IN || g->weight_type != MI355G_W_F32)
      return fail(MI355G_EINVAL, "RIPSCSR1 holds incoming-orientation fp32 graphs");
    File f(path, "wb");
    if (!f.f) return fail(MI355G_EIO, std::string("cannot create ") + path);
    const uint64_t nnz = (uint64_t)g->nnz();
    const uint64_t h[10] = {1, 1, (uint64_t)g->rows, (uint64_t)g->cols, g->declared_edges,
                            g->loaded_edges, nnz, (uint64_t)g->rowptr.size(), nnz, nnz};
    if (!write_arr(f.f, "RIPSCSR1", 8) || !write_arr(f.f, h, 10) ||
        !write_arr(f.f, g->rowptr.data(), g->rowptr.size()) ||
        !write_arr(f.f, g->col.data(), g->col.size()) ||
        !write_arr(f.f, g->w.data(), g->w.size()))
      return fail(MI355G_EIO, "short write");
    return 0;
  });
}

// ---- RIPSADJ1 (dense incoming matrix) ---------------------------------------
int mi355g_write_ripsadj1(const mi355g_graph* g, const char* path) {
  return guarded([&] {
    if (!g || g->orientation != MI355G_IN || g->weight_type != MI355G_W_F32)
      return fail(MI355G_EINVAL, "RIPSADJ1 holds incoming-orientation fp32 graphs");
    const uint64_t n = (uint64_t)g->rows;
    std::vector<float> a(n * n, std::numeric_limits<float>::infinity());
    for (uint64_t v = 0; v < n; ++v)
      for (int64_t e = g->rowptr[v]; e < g->rowptr[v + 1]; ++e) {
        float x;
        std::memcpy(&x, &g->w[e], 4);
        float& slot = a[v * n + (uint64_t)g->col[e]];
        if (x < slot) slot = x;
      }
    File f(path, "wb");
    if (!f.f) return fail(MI355G_EIO, std::string("cannot create ") + path);
    const uint64_t h[6] = {1, 1, n, g->declared_edges, g->loaded_edges, n * n};
    if (!write_arr(f.f, "RIPSADJ1", 8) || !write_arr(f.f, h, 6) || !write_arr(f.f, a.data(), a.size()))
      return fail(MI355G_EIO, "short write");
    return 0;
  });
}

int mi355g_read_ripsadj1(const char* path, float* matrix, int64_t capacity, int64_t* n_out) {
  return guarded([&] {
    File f(path, "rb");
    if (!f.f) return fail(MI355G_EIO, std::string("cannot open ") + path);
    char magic[8];
    uint64_t h[6];
    if (!read_arr(f.f, magic, 8) || std::memcmp(magic, "RIPSADJ1", 8) != 0)
      return fail(MI355G_EFORMAT, "bad RIPSADJ1 magic");
    if (!read_arr(f.f, h, 6)) return fail(MI355G_EFORMAT, "truncated RIPSADJ1 header");
    if (h[0] != 1 || h[1] != 1 || h[5] != h[2] * h[2])
      return fail(MI355G_EFORMAT, "unsupported RIPSADJ1 header");
    if (n_out) *n_out = (int64_t)h[2];
    if (!matrix) return 0;
    if ((uint64_t)capacity < h[5]) return fail(MI355G_EINVAL, "matrix buffer too small");
    if (!read_arr(f.f, matrix, h[5])) return fail(MI355G_EFORMAT, "truncated RIPSADJ1 matrix");
    return 0;
  });
}

// Dense row-major matrix -> CSR of its finite entries.  With
// orientation MI355G_OUT the matrix is adj[u*n+v] = w(u->v) and the result is
// outgoing (ONE/sssp_hip.cpp:320-348: the diagonal and non-finite entries
// are skipped, negative weights rejected); with MI355G_IN it is the
// incoming matrix A[v*n+u] = w(u->v) (BF/src/bellman_ford_hip.hpp:8-59).
int mi355g_from_dense(const float* adj, int64_t n, int32_t orientation, int32_t skip_diagonal,
                      int32_t reject_negative, mi355g_graph** out) {
  return guarded([&] {
    if (!adj || n <= 0 || !out) return fail(MI355G_EINVAL, "invalid dense arguments");
    auto g = new mi355g_graph();
    g->rows = g->cols = n;
    g->weight_type = MI355G_W_F32;
    g->orientation = orientation;
    g->rowptr.assign(n + 1, 0);
    for (int64_t r = 0; r < n; ++r) {
      for (int64_t c = 0; c < n; ++c) {
        const float x = adj[r * n + c];
        if ((skip_diagonal && r == c) || !std::isfinite(x)) continue;
        if (reject_negative && x < 0.0f) {
          delete g;
          return fail(MI355G_EINVAL, "negative edge weight in dense adjacency");
        }
        uint32_t b;
        std::memcpy(&b, &x, 4);
        g->col.push_back((int32_t)c);
        g->w.push_back(b);
      }
      g->rowptr[r + 1] = (int64_t)g->col.size();
    }
    g->declared_edges = g->loaded_edges = (uint64_t)g->col.size();
    *out = g;
    return 0;
  });
}

// ---- DIMACS text -----------------------------------------------------------
int mi355g_read_dimacs(const char* path, mi355g_graph** out) {
  return guarded([&] {
    File f(path, "r");
    if (!f.f) return fail(MI355G_EIO, std::string("cannot open ") + path);
    long long n = -1, declared = 0;
    std::vector<int32_t> src, dst;
    std::vector<uint32_t> wv;
    char line[512];
    long long lineno = 0;
    while (std::fgets(line, sizeof line, f.f)) {
      ++lineno;
      if (line[0] == 'c' || line[0] == '\n' || line[0] == '\r' || line[0] == 0) continue;
      if (line[0] == 'p') {
        char prob[32];
        long long nn, mm;
        if (std::sscanf(line, "p %31s %lld %lld", prob, &nn, &mm) != 3 || nn <= 0 ||
            nn > std::numeric_limits<int32_t>::max() || mm < 0)
          return fail(MI355G_EFORMAT, "invalid problem line " + std::to_string(lineno));
        n = nn;
        declared = mm;
        continue;
      }
      if (line[0] == 'a') {
        if (n < 0) return fail(MI355G_EFORMAT, "edge line before problem line");
        long long a, b;
        double wt;
        if (std::sscanf(line, "a %lld %lld %lf", &a, &b, &wt) != 3)
          return fail(MI355G_EFORMAT, "invalid edge line " + std::to_string(lineno));
        if (a <= 0 || a > n || b <= 0 || b > n)
          return fail(MI355G_EFORMAT, "edge endpoint outside graph at line " +
                                          std::to_string(lineno));
        if (!std::isfinite(wt))
          return fail(MI355G_EFORMAT, "edge weight not finite at line " + std::to_string(lineno));
        const float fw = (float)wt;
        uint32_t bits;
        std::memcpy(&bits, &fw, 4);
        src.push_back((int32_t)(a - 1));
        dst.push_back((int32_t)(b - 1));
        wv.push_back(bits);
      }
    }
    if (n < 0) return fail(MI355G_EFORMAT, "missing DIMACS problem line");
    auto g = build_from_edges(n, (int64_t)src.size(), src.data(), dst.data(), wv.data(),
                              MI355G_W_F32, MI355G_IN, true);
    g->declared_edges = (uint64_t)declared;
    g->loaded_edges = (uint64_t)src.size();
    *out = g;
    return 0;
  });
}

// ---- Galois .gr v1 ---------------------------------------------------------
int mi355g_read_galois_gr(const char* path, mi355g_graph** out) {
  return guarded([&] {
    File f(path, "rb");
    if (!f.f) return fail(MI355G_EIO, std::string("cannot open ") + path);
    uint64_t h[4];
    if (!read_arr(f.f, h, 4)) return fail(MI355G_EFORMAT, "file too small to be a .gr graph");
    if (h[0] != 1) return fail(MI355G_EFORMAT, "unsupported .gr version");
    if (h[1] != 0 && h[1] != 4) return fail(MI355G_EFORMAT, "edge data must be absent or 4 bytes");
    if (h[2] == 0 || h[2] > (uint64_t)std::numeric_limits<int32_t>::max() ||
        h[3] > (uint64_t)std::numeric_limits<int32_t>::max())
      return fail(MI355G_EFORMAT, ".gr graph exceeds 32-bit limits");
    const uint64_t n = h[2], m = h[3];
    // the tables must be in the file before anything is allocated for them
    const int64_t size = file_bytes(f.f);
    if (size < 0) return fail(MI355G_EIO, "cannot size the file");
    if (32 + 8 * n + 4 * m > (uint64_t)size) return fail(MI355G_EFORMAT, "truncated .gr tables");
    std::vector<uint64_t> ends(n);
    std::vector<uint32_t> dsts(m);
    if (!read_arr(f.f, ends.data(), n)) return fail(MI355G_EFORMAT, "truncated row-offset table");
    if (!read_arr(f.f, dsts.data(), m)) return fail(MI355G_EFORMAT, "truncated destination table");
    std::vector<uint32_t> wbits(m);
    if (h[1] == 4) {
      if (m & 1) {
        uint32_t pad;
        if (!read_arr(f.f, &pad, 1)) return fail(MI355G_EFORMAT, "truncated padding");
      }
      if (!read_arr(f.f, wbits.data(), m)) return fail(MI355G_EFORMAT, "truncated edge-weight table");
    } else {
      const float one = 1.0f;
      uint32_t b;
      std::memcpy(&b, &one, 4);
      std::fill(wbits.begin(), wbits.end(), b);
    }
    auto g = new mi355g_graph();
    g->rows = g->cols = (int64_t)n;
    g->weight_type = MI355G_W_F32;
    g->orientation = MI355G_OUT;
    g->declared_edges = g->loaded_edges = m;
    g->rowptr.assign(n + 1, 0);
    for (uint64_t v = 0; v < n; ++v) {
      if (ends[v] > m || (int64_t)ends[v] < g->rowptr[v]) {
        delete g;
        return fail(MI355G_EFORMAT, "invalid row offset for node " + std::to_string(v));
      }
      g->rowptr[v + 1] = (int64_t)ends[v];
    }
    if ((uint64_t)g->rowptr[n] != m) {
      delete g;
      return fail(MI355G_EFORMAT, "final row offset differs from the header edge count");
    }
    g->col.resize(m);
    for (uint64_t e = 0; e < m; ++e) {
      if (dsts[e] >= n) {
        delete g;
        return fail(MI355G_EFORMAT, "edge destination out of range");
      }
      float fw;
      std::memcpy(&fw, &wbits[e], 4);
      if (!std::isfinite(fw) || fw < 0.0f) {
        delete g;
        return fail(MI355G_EFORMAT, "SSSP requires finite nonnegative weights");
      }
      g->col[e] = (int32_t)dsts[e];
    }
    g->w = std::move(wbits);
    *out = g;
    return 0;
  });
}

int mi355g_write_galois_gr(const mi355g_graph* g, const char* path) {
  return guarded([&] {
    if (!g || g->orientation != MI355G_OUT || g->weight_type != MI355G_W_F32)
      return fail(MI355G_EINVAL, ".gr holds outgoing-orientation fp32 graphs");
    File f(path, "wb");
    if (!f.f) return fail(MI355G_EIO, std::string("cannot create ") + path);
    const uint64_t m = (uint64_t)g->nnz();
    const uint64_t h[4] = {1, 4, (uint64_t)g->rows, m};
    std::vector<uint64_t> ends(g->rows);
    for (int64_t v = 0; v < g->rows; ++v) ends[v] = (uint64_t)g->rowptr[v + 1];
    std::vector<uint32_t> d(g->col.begin(), g->col.end());
    const uint32_t pad = 0;
    if (!write_arr(f.f, h, 4) || !write_arr(f.f, ends.data(), ends.size()) ||
        !write_arr(f.f, d.data(), d.size()) || ((m & 1) && !write_arr(f.f, &pad, 1)) ||
        !write_arr(f.f, g->w.data(), g->w.size()))
      return fail(MI355G_EIO, "short write");
    return 0;
  });
}

// ---- generators --------------------------------------------------------------
int mi355g_gen_grid4(int32_t rows, int32_t cols, uint32_t wmin, uint32_t wmax, uint64_t seed,
                     mi355g_graph** out) {
  return guarded([&] {
    if (!out || rows <= 0 || cols <= 0 || wmin > wmax)
      return fail(MI355G_EINVAL, "invalid grid arguments");
    const int64_t n = (int64_t)rows * cols;
    auto g = new mi355g_graph();
    g->rows = g->cols = n;
    g->weight_type = MI355G_W_U32;
    g->orientation = MI355G_OUT;
    g->rowptr.assign(n + 1, 0);
    for (int64_t u = 0; u < n; ++u) {
      const int32_t r = (int32_t)(u / cols), c = (int32_t)(u % cols);
      g->rowptr[u + 1] = g->rowptr[u] + (r > 0) + (c > 0) + (c + 1 < cols) + (r + 1 < rows);
    }
    const int64_t m = g->rowptr[n];
    g->col.resize(m);
    g->w.resize(m);
    const uint64_t range = (uint64_t)wmax - wmin + 1;
    for (int64_t u = 0; u < n; ++u) {
      const int32_t r = (int32_t)(u / cols), c = (int32_t)(u % cols);
      int64_t p = g->rowptr[u];
      // directions in ascending destination order: up, left, right, down
      const int64_t nb[4] = {r > 0 ? u - cols : -1, c > 0 ? u - 1 : -1,
                             c + 1 < cols ? u + 1 : -1, r + 1 < rows ? u + cols : -1};
      for (int d = 0; d < 4; ++d) {
        if (nb[d] < 0) continue;
        g->col[p] = (int32_t)nb[d];
        g->w[p] = wmin + (uint32_t)range_u64(rng_u64(seed, 30, (uint64_t)u * 4 + d), range);
        ++p;
      }
    }
    g->declared_edges = g->loaded_edges = (uint64_t)m;
    *out = g;
    return 0;
  });
}

int mi355g_gen_rmat(int32_t scale, int32_t edgefactor, double a, double b, double c,
                    uint64_t seed, int32_t permute, mi355g_graph** out) {
  return guarded([&] {
    if (!out || scale < 1 || scale > 30 || edgefactor < 1 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0)
      return fail(MI355G_EINVAL, "invalid R-MAT arguments");
    const int64_t n = (int64_t)1 << scale;
    const int64_t draws = (int64_t)edgefactor << scale;
    const uint64_t ta = (uint64_t)(a * 4294967296.0);
    const uint64_t tb = (uint64_t)((a + b) * 4294967296.0);
    const uint64_t tc = (uint64_t)((a + b + c) * 4294967296.0);
    // Seeded permutation (Fisher-Yates driven by the counter stream).
    std::vector<int32_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    if (permute) {
      for (int64_t i = n - 1; i > 0; --i) {
        const int64_t j = (int64_t)range_u64(rng_u64(seed, 2, (uint64_t)i), (uint64_t)i + 1);
        std::swap(perm[i], perm[j]);
      }
    }
    // Draw edges; key = lo * n + hi for the undirected pair, self loops dropped.
    std::vector<uint64_t> keys(draws);
    std::vector<uint8_t> keep(draws);
    parallel_for(draws, [&](int64_t lo, int64_t hi, unsigned) {
      for (int64_t i = lo; i < hi; ++i) {
        uint64_t u = 0, v = 0;
        uint64_t x = 0;
        for (int l = 0; l < scale; ++l) {
          uint64_t r32;
          if ((l & 1) == 0) {
            x = rng_u64(seed, 1, (uint64_t)i * 16 + (uint64_t)(l >> 1));
            r32 = x >> 32;
          } else {
            r32 = x & 0xffffffffull;
          }
          const int q = r32 < ta ? 0 : r32 < tb ? 1 : r32 < tc ? 2 : 3;
          u = (u << 1) | (uint64_t)(q >> 1);
          v = (v << 1) | (uint64_t)(q & 1);
        }
        const uint64_t pu = (uint64_t)perm[u], pv = (uint64_t)perm[v];
        keep[i] = pu != pv;
        const uint64_t lo_v = std::min(pu, pv), hi_v = std::max(pu, pv);
        keys[i] = lo_v * (uint64_t)n + hi_v;
      }
    });
    {
      int64_t k = 0;
      for (int64_t i = 0; i < draws; ++i)
        if (keep[i]) keys[k++] = keys[i];
      keys.resize(k);
      std::vector<uint8_t>().swap(keep);
    }
    radix_sort_u64(keys, 2 * scale);
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int64_t und = (int64_t)keys.size();
    auto g = new mi355g_graph();
    g->rows = g->cols = n;
    g->weight_type = MI355G_W_F32;
    g->orientation = MI355G_OUT;
    g->rowptr.assign(n + 1, 0);
    for (int64_t i = 0; i < und; ++i) {
      g->rowptr[keys[i] / n + 1]++;
      g->rowptr[keys[i] % n + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) g->rowptr[i + 1] += g->rowptr[i];
    g->col.resize(2 * und);
    g->w.resize(2 * und);
    std::vector<int64_t> cur(g->rowptr.begin(), g->rowptr.end() - 1);
    // Ascending (lo, hi) order fills every row in ascending column order:
    // neighbours below a vertex arrive (as hi) before those above it (as lo).
    for (int64_t i = 0; i < und; ++i) {
      const int64_t lo = (int64_t)(keys[i] / n), hi = (int64_t)(keys[i] % n);
      const float wf = unit_f32(rng_u64(seed, 3, keys[i]));
      uint32_t bits;
      std::memcpy(&bits, &wf, 4);
      int64_t p = cur[lo]++;
      g->col[p] = (int32_t)hi;
      g->w[p] = bits;
      p = cur[hi]++;
      g->col[p] = (int32_t)lo;
      g->w[p] = bits;
    }
    g->declared_edges = (uint64_t)draws;
    g->loaded_edges = (uint64_t)(2 * und);
    *out = g;
    return 0;
  });
}

int mi355g_gen_routing_grid(int32_t side, double p_delete, double p_diag, double mu,
                            double sigma, uint64_t seed, mi355g_graph** out) {
  return guarded([&] {
    if (!out || side < 2 || p_delete < 0 || p_delete > 1 || p_diag < 0 || p_diag > 1)
      return fail(MI355G_EINVAL, "invalid routing-grid arguments");
    const int64_t S = side, n = S * S;
    const uint64_t tdel = p_delete >= 1.0 ? ~0ull : (uint64_t)(p_delete * 18446744073709551616.0);
    const uint64_t tdiag = p_diag >= 1.0 ? ~0ull : (uint64_t)(p_diag * 18446744073709551616.0);
    // Undirected edge ids per kind (stream): 10 horizontal (r*(S-1)+c),
    // 11 vertical (r*S+c), 12 diagonal down-right of cell (r*(S-1)+c),
    // 13 diagonal down-left of cell (r*(S-1)+c) joining (r,c+1)-(r+1,c).
    auto h_exists = [&](int64_t r, int64_t c) {  // (r,c)-(r,c+1)
      return rng_u64(seed, 10, (uint64_t)(r * (S - 1) + c)) >= tdel;
    };
    auto v_exists = [&](int64_t r, int64_t c) {  // (r,c)-(r+1,c)
      return rng_u64(seed, 11, (uint64_t)(r * S + c)) >= tdel;
    };
    auto d1_exists = [&](int64_t r, int64_t c) {  // (r,c)-(r+1,c+1)
      return rng_u64(seed, 12, (uint64_t)(r * (S - 1) + c)) < tdiag;
    };
    auto d2_exists = [&](int64_t r, int64_t c) {  // (r,c+1)-(r+1,c)
      return rng_u64(seed, 13, (uint64_t)(r * (S - 1) + c)) < tdiag;
    };
    auto weight = [&](uint64_t kind, uint64_t id) -> uint32_t {
      const uint64_t x = rng_u64(seed, 20 + kind, id);
      const double u1 = open_unit_f64((uint32_t)(x >> 32));
      const double u2 = open_unit_f64((uint32_t)x);
      const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      double wv = std::ceil(std::exp(mu + sigma * z));
      if (wv < 1.0) wv = 1.0;
      if (wv > 2147483647.0) wv = 2147483647.0;
      return (uint32_t)wv;
    };
    // For vertex (r,c) enumerate the 8 potential neighbours in ascending id
    // order, returning (neighbour, kind, id) for those that exist.
    struct Nb {
      int64_t v;
      uint64_t kind, id;
    };
    auto neighbours = [&](int64_t r, int64_t c, Nb* nb) {
      int k = 0;
      if (r > 0) {
        if (c > 0 && d1_exists(r - 1, c - 1)) nb[k++] = {(r - 1) * S + c - 1, 2, (uint64_t)((r - 1) * (S - 1) + c - 1)};
        if (v_exists(r - 1, c)) nb[k++] = {(r - 1) * S + c, 1, (uint64_t)((r - 1) * S + c)};
        if (c + 1 < S && d2_exists(r - 1, c)) nb[k++] = {(r - 1) * S + c + 1, 3, (uint64_t)((r - 1) * (S - 1) + c)};
      }
      if (c > 0 && h_exists(r, c - 1)) nb[k++] = {r * S + c - 1, 0, (uint64_t)(r * (S - 1) + c - 1)};
      if (c + 1 < S && h_exists(r, c)) nb[k++] = {r * S + c + 1, 0, (uint64_t)(r * (S - 1) + c)};
      if (r + 1 < S) {
        if (c > 0 && d2_exists(r, c - 1)) nb[k++] = {(r + 1) * S + c - 1, 3, (uint64_t)(r * (S - 1) + c - 1)};
        if (v_exists(r, c)) nb[k++] = {(r + 1) * S + c, 1, (uint64_t)(r * S + c)};
        if (c + 1 < S && d1_exists(r, c)) nb[k++] = {(r + 1) * S + c + 1, 2, (uint64_t)(r * (S - 1) + c)};
      }
      return k;
    };
    auto g = new mi355g_graph();
    g->rows = g->cols = n;
    g->weight_type = MI355G_W_U32;
    g->orientation = MI355G_OUT;
    g->rowptr.assign(n + 1, 0);
    parallel_for(S, [&](int64_t r0, int64_t r1, unsigned) {
      Nb nb[8];
      for (int64_t r = r0; r < r1; ++r)
        for (int64_t c = 0; c < S; ++c) g->rowptr[r * S + c + 1] = neighbours(r, c, nb);
    });
    for (int64_t i = 0; i < n; ++i) g->rowptr[i + 1] += g->rowptr[i];
    const int64_t m = g->rowptr[n];
    g->col.resize(m);
    g->w.resize(m);
    parallel_for(S, [&](int64_t r0, int64_t r1, unsigned) {
      Nb nb[8];
      for (int64_t r = r0; r < r1; ++r)
        for (int64_t c = 0; c < S; ++c) {
          const int k = neighbours(r, c, nb);
          int64_t p = g->rowptr[r * S + c];
          for (int i = 0; i < k; ++i, ++p) {
            g->col[p] = (int32_t)nb[i].v;
            g->w[p] = weight(nb[i].kind, nb[i].id);
          }
        }
    });
    g->declared_edges = g->loaded_edges = (uint64_t)m;
    *out = g;
    return 0;
  });
}

int mi355g_gen_chain_random(int64_t n, int32_t avg_deg, double wmin, double wmax,
                            int32_t weight_type, uint64_t seed, mi355g_graph** out) {
  return guarded([&] {
    if (!out || n <= 1 || avg_deg < 0 || !(wmin >= 0) || !(wmax >= wmin))
      return fail(MI355G_EINVAL, "invalid chain-random arguments");
    const int64_t m = (n - 1) + n * (int64_t)avg_deg;
    std::vector<int32_t> s(m), d(m);
    std::vector<uint32_t> wv(m);
    int64_t k = 0;
    auto wbits = [&](uint64_t x) -> uint32_t {
      if (weight_type == MI355G_W_U32) {
        const uint64_t lo = (uint64_t)wmin, hi = (uint64_t)wmax;
        return (uint32_t)(lo + range_u64(x, hi - lo + 1));
      }
      const float f = (float)(wmin + (wmax - wmin) * ((double)(x >> 11) * (1.0 / 9007199254740992.0)));
      uint32_t b;
      std::memcpy(&b, &f, 4);
      return b;
    };
    const float one = 1.0f;
    uint32_t one_bits;
    std::memcpy(&one_bits, &one, 4);
    for (int64_t u = 0; u + 1 < n; ++u, ++k) {
      s[k] = (int32_t)u;
      d[k] = (int32_t)(u + 1);
      wv[k] = weight_type == MI355G_W_U32 ? 1u : one_bits;
    }
    for (int64_t u = 0; u < n; ++u)
      for (int32_t j = 0; j < avg_deg; ++j, ++k) {
        const uint64_t idx = (uint64_t)u * (uint64_t)avg_deg + (uint64_t)j;
        int64_t v = (int64_t)range_u64(rng_u64(seed, 50, idx), (uint64_t)n);
        if (v == u) v = (v + 1) % n;
        s[k] = (int32_t)u;
        d[k] = (int32_t)v;
        wv[k] = wbits(rng_u64(seed, 51, idx));
      }
    *out = build_from_edges(n, m, s.data(), d.data(), wv.data(), weight_type, MI355G_OUT, false);
    return 0;
  });
}

int mi355g_pick_sources(const mi355g_graph* g, int32_t count, uint64_t seed, int32_t* sources) {
  return guarded([&] {
    if (!g || count < 0 || (count > 0 && !sources)) return fail(MI355G_EINVAL, "invalid arguments");
    if (g->orientation != MI355G_OUT) return fail(MI355G_EINVAL, "sources need an outgoing graph");
    int32_t found = 0;
    for (uint64_t j = 0; found < count && j < (uint64_t)count * 100000ull + 1000000ull; ++j) {
      const int64_t v = (int64_t)range_u64(rng_u64(seed, 40, j), (uint64_t)g->rows);
      if (g->rowptr[v + 1] == g->rowptr[v]) continue;
      bool dup = false;
      for (int32_t i = 0; i < found; ++i) dup |= sources[i] == v;
      if (!dup) sources[found++] = (int32_t)v;
    }
    if (found < count) return fail(MI355G_EINVAL, "not enough vertices with out-degree > 0");
    return 0;
  });
}

int mi355g_sha256(const mi355g_graph* g, char* hex65) {
  return guarded([&] {
    if (!g || !hex65) return fail(MI355G_EINVAL, "null argument");
    Sha256 s;
    s.update(g->rowptr.data(), g->rowptr.size() * 8);
    s.update(g->col.data(), g->col.size() * 4);
    s.update(g->w.data(), g->w.size() * 4);
    s.final_hex(hex65);
    hex65[64] = 0;
    return 0;
  });
}

}  // extern "C"

// ---- RIPSIFM1 ------------------------------------------------------------
// Layout as written by Routing/interchange_to_csr.cpp:1390-1587 (all u64
// little-endian): magic "RIPSIFM1"; version, orientation; the 13 counts and
// 4 path / design-name string references of mi355g_ifm_header_t; the strings
// (u64 length + bytes); node device ids; (tile, pip) per edge; (wire0, wire1,
// forward) per pip; (node, site, pin) per sink site pin; per route request
// net string, logical net index, source count + (node, site, pin) records,
// sink count + records; logical cells (3 u64), nets (4), port instances (7);
// blocked nodes; sink-stop nodes; the physical and logical netlist bytes.
namespace {

constexpr uint64_t kIfmNoIndex = ~0ull;

struct IfmReader {
  FILE* f;
  uint64_t left;  // bytes not yet read
  bool u64(uint64_t& v) {
    if (left < 8 || std::fread(&v, 8, 1, f) != 1) return false;
    left -= 8;
    return true;
  }
  // n records of k u64 each; refuses counts the file cannot hold before
  // allocating
  bool table(std::vector<uint64_t>& out, uint64_t n, uint64_t k) {
    if (k && n > left / (8 * k)) return false;
    out.resize((size_t)(n * k));
    if (!read_arr(f, out.data(), out.size())) return false;
    left -= n * k * 8;
    return true;
  }
  bool bytes(std::vector<uint8_t>& out, uint64_t n) {
    if (n > left) return false;
    out.resize((size_t)n);
    if (!read_arr(f, out.data(), out.size())) return false;
    left -= n;
    return true;
  }
};

}  // namespace

extern "C" {

int mi355g_read_ripsifm1(const char* path, mi355g_ifm** out) {
  return guarded([&] {
    if (!path || !out) return fail(MI355G_EINVAL, "null argument");
    File f(path, "rb");
    if (!f.f) return fail(MI355G_EIO, std::string("cannot open ") + path);
    if (std::fseek(f.f, 0, SEEK_END) != 0) return fail(MI355G_EIO, "cannot size the file");
    const long size = std::ftell(f.f);
    std::rewind(f.f);
    if (size < 8) return fail(MI355G_EFORMAT, "truncated RIPSIFM1 file");
    IfmReader r{f.f, (uint64_t)size};
    char magic[8];
    if (!read_arr(f.f, magic, 8) || std::memcmp(magic, "RIPSIFM1", 8) != 0)
      return fail(MI355G_EFORMAT, "bad RIPSIFM1 magic");
    r.left -= 8;
    auto m = std::make_unique<mi355g_ifm>();
    mi355g_ifm_header_t& h = m->h;
    uint64_t* fields[] = {&h.version, &h.orientation, &h.string_count, &h.node_count,
                          &h.edge_attr_count, &h.pip_data_count, &h.site_pin_attr_count,
                          &h.route_request_count, &h.blocked_node_count, &h.sink_stop_node_count,
                          &h.logical_cell_count, &h.logical_net_count,
                          &h.logical_port_instance_count, &h.physical_netlist_byte_count,
                          &h.logical_netlist_byte_count, &h.device_path_string,
                          &h.physical_path_string, &h.logical_path_string,
                          &h.logical_design_name_string};
    for (uint64_t* v : fields)
      if (!r.u64(*v)) return fail(MI355G_EFORMAT, "truncated RIPSIFM1 header");
    if (h.version != 2) return fail(MI355G_EFORMAT, "unsupported RIPSIFM1 version (expected 2)");
    if (h.orientation != 1) return fail(MI355G_EFORMAT, "RIPSIFM1 orientation must be 1 (incoming)");
    // strings
    if (h.string_count > r.left / 8) return fail(MI355G_EFORMAT, "truncated RIPSIFM1 strings");
    m->str_off.assign((size_t)h.string_count + 1, 0);
    for (uint64_t i = 0; i < h.string_count; ++i) {
      uint64_t len = 0;
      if (!r.u64(len) || len > r.left) return fail(MI355G_EFORMAT, "truncated RIPSIFM1 strings");
      const size_t at = m->str_bytes.size();
      m->str_bytes.resize(at + (size_t)len);
      if (!read_arr(f.f, m->str_bytes.data() + at, (size_t)len))
        return fail(MI355G_EFORMAT, "truncated RIPSIFM1 strings");
      r.left -= len;
      m->str_off[(size_t)i + 1] = m->str_bytes.size();
    }
    if (!r.table(m->node_ids, h.node_count, 1) || !r.table(m->edge_attrs, h.edge_attr_count, 2) ||
        !r.table(m->pips, h.pip_data_count, 3) || !r.table(m->site_pins, h.site_pin_attr_count, 3))
      return fail(MI355G_EFORMAT, "truncated RIPSIFM1 tables");
    // route requests: variable-length source / sink lists
    m->nets.reserve((size_t)std::min<uint64_t>(h.route_request_count, r.left / 32) * 2);
    m->src_off.assign(1, 0);
    m->sink_off.assign(1, 0);
    for (uint64_t q = 0; q < h.route_request_count; ++q) {
      uint64_t net = 0, lnet = 0, k = 0;
      if (!r.u64(net) || !r.u64(lnet) || !r.u64(k) || k > r.left / 24)
        return fail(MI355G_EFORMAT, "truncated RIPSIFM1 route requests");
      m->nets.push_back(net);
      m->nets.push_back(lnet);
      std::vector<uint64_t> rec;
      if (!r.table(rec, k, 3)) return fail(MI355G_EFORMAT, "truncated RIPSIFM1 route requests");
      m->sources.insert(m->sources.end(), rec.begin(), rec.end());
      m->src_off.push_back(m->sources.size() / 3);
      if (!r.u64(k) || !r.table(rec, k, 3)) return fail(MI355G_EFORMAT, "truncated RIPSIFM1 route requests");
      m->sinks.insert(m->sinks.end(), rec.begin(), rec.end());
      m->sink_off.push_back(m->sinks.size() / 3);
    }
    h.route_source_count = m->sources.size() / 3;
    h.route_sink_count = m->sinks.size() / 3;
    if (!r.table(m->cells, h.logical_cell_count, 3) || !r.table(m->lnets, h.logical_net_count, 4) ||
        !r.table(m->ports, h.logical_port_instance_count, 7) ||
        !r.table(m->blocked, h.blocked_node_count, 1) || !r.table(m->sink_stop, h.sink_stop_node_count, 1) ||
        !r.bytes(m->physical, h.physical_netlist_byte_count) ||
        !r.bytes(m->logical, h.logical_netlist_byte_count))
      return fail(MI355G_EFORMAT, "truncated RIPSIFM1 trailing sections");
    if (r.left != 0) return fail(MI355G_EFORMAT, "trailing bytes after the RIPSIFM1 sections");
    // references
    auto str_ok = [&](uint64_t s, bool none_ok) {
      return s < h.string_count || (none_ok && s == kIfmNoIndex);
    };
    for (uint64_t i = 0; i < h.edge_attr_count; ++i)
      if (!str_ok(m->edge_attrs[2 * i], false) || m->edge_attrs[2 * i + 1] >= h.pip_data_count)
        return fail(MI355G_EFORMAT, "RIPSIFM1 edge attribute refers outside the tables");
    for (uint64_t i = 0; i < h.pip_data_count; ++i)
      if (!str_ok(m->pips[3 * i], false) || !str_ok(m->pips[3 * i + 1], false) || m->pips[3 * i + 2] > 1)
        return fail(MI355G_EFORMAT, "RIPSIFM1 pip record refers outside the tables");
    auto pins_ok = [&](const std::vector<uint64_t>& v) {
      for (size_t i = 0; i < v.size(); i += 3)
        if (v[i] >= h.node_count || !str_ok(v[i + 1], false) || !str_ok(v[i + 2], false)) return false;
      return true;
    };
    if (!pins_ok(m->site_pins) || !pins_ok(m->sources) || !pins_ok(m->sinks))
      return fail(MI355G_EFORMAT, "RIPSIFM1 site pin refers outside the tables");
    for (size_t q = 0; q < m->nets.size(); q += 2)
      if (!str_ok(m->nets[q], false) ||
          (m->nets[q + 1] != kIfmNoIndex && m->nets[q + 1] >= h.logical_net_count))
        return fail(MI355G_EFORMAT, "RIPSIFM1 route request refers outside the tables");
    for (uint64_t v : m->blocked)
      if (v >= h.node_count) return fail(MI355G_EFORMAT, "RIPSIFM1 blocked node out of range");
    for (uint64_t v : m->sink_stop)
      if (v >= h.node_count) return fail(MI355G_EFORMAT, "RIPSIFM1 sink-stop node out of range");
    for (uint64_t s2 : {h.device_path_string, h.physical_path_string, h.logical_path_string,
                        h.logical_design_name_string})
      if (!str_ok(s2, true)) return fail(MI355G_EFORMAT, "RIPSIFM1 header string out of range");
    *out = m.release();
    return 0;
  });
}

void mi355g_ifm_free(mi355g_ifm* m) { delete m; }

int mi355g_ifm_header(const mi355g_ifm* m, mi355g_ifm_header_t* out) {
  if (!m || !out) return fail(MI355G_EINVAL, "null argument");
  *out = m->h;
  return 0;
}

int mi355g_ifm_string(const mi355g_ifm* m, uint64_t index, const char** data, uint64_t* len) {
  if (!m || !data || !len) return fail(MI355G_EINVAL, "null argument");
  if (index >= m->h.string_count) return fail(MI355G_EINVAL, "string index out of range");
  *data = m->str_bytes.data() + m->str_off[(size_t)index];
  *len = m->str_off[(size_t)index + 1] - m->str_off[(size_t)index];
  return 0;
}

namespace {
void copy_out(const std::vector<uint64_t>& v, uint64_t* dst) {
  if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * 8);
}
}  // namespace

int mi355g_ifm_tables(const mi355g_ifm* m, uint64_t* node_device_ids, uint64_t* edge_attrs,
                      uint64_t* pip_data, uint64_t* site_pins, uint64_t* blocked,
                      uint64_t* sink_stop) {
  if (!m) return fail(MI355G_EINVAL, "null argument");
  copy_out(m->node_ids, node_device_ids);
  copy_out(m->edge_attrs, edge_attrs);
  copy_out(m->pips, pip_data);
  copy_out(m->site_pins, site_pins);
  copy_out(m->blocked, blocked);
  copy_out(m->sink_stop, sink_stop);
  return 0;
}

int mi355g_ifm_requests(const mi355g_ifm* m, uint64_t* nets, uint64_t* src_offsets,
                        uint64_t* sources, uint64_t* sink_offsets, uint64_t* sinks) {
  if (!m) return fail(MI355G_EINVAL, "null argument");
  copy_out(m->nets, nets);
  copy_out(m->src_off, src_offsets);
  copy_out(m->sources, sources);
  copy_out(m->sink_off, sink_offsets);
  copy_out(m->sinks, sinks);
  return 0;
}

int mi355g_ifm_logical(const mi355g_ifm* m, uint64_t* cells, uint64_t* nets, uint64_t* ports) {
  if (!m) return fail(MI355G_EINVAL, "null argument");
  copy_out(m->cells, cells);
  copy_out(m->lnets, nets);
  copy_out(m->ports, ports);
  return 0;
}

int mi355g_ifm_payloads(const mi355g_ifm* m, const uint8_t** physical, const uint8_t** logical) {
  if (!m) return fail(MI355G_EINVAL, "null argument");
  if (physical) *physical = m->physical.data();
  if (logical) *logical = m->logical.data();
  return 0;
}

}  // extern "C"
