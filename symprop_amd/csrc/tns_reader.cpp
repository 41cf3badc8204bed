// .tns input: the reference's text format read in parallel, and a binary
// tuple cache.
//
// Text semantics (src/utils/loader.cpp:8-36): a header "order dim", then
// whitespace-separated tokens read `order` indices and one value at a time;
// each tuple is sorted; reading stops at the first failed read, and a tuple
// whose value token ends the file (no whitespace after it) is dropped, since
// the stream's good() is then false.
//
// Fast path: the file is read once, split into chunks at whitespace, tokens
// are counted per chunk (parallel), a prefix sum gives every chunk its first
// token's place, and the tokens are parsed in place (parallel).  A file the
// fast path cannot reproduce exactly -- a token that is not a plain unsigned
// integer or decimal number, an index above 2^32-1, a value starting with a
// sign '+' or spelled inf/nan -- is read again by the serial stream reader,
// whose result is the reference's by construction.
#include "sp_internal.hpp"

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <thread>

namespace spamd {

namespace {

struct TnsData {
  u32 order = 0, dim = 0;
  std::vector<u32> inds;
  std::vector<double> vals;
};

inline bool is_ws(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// Serial reader with the stream semantics.
bool read_serial(const char *path, TnsData &d) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  u32 o = 0, dm = 0;
  f >> o >> dm;
  d.order = o;
  d.dim = dm;
  d.inds.clear();
  d.vals.clear();
  if (o == 0 || o > 64) return true;
  std::vector<u32> t(o);
  while (true) {
    for (u32 i = 0; i < o; ++i) {
      u32 v = 0;
      f >> v;
      t[i] = v;
    }
    double v = 0;
    f >> v;
    if (!f.good()) break;
    std::sort(t.begin(), t.end());
    d.inds.insert(d.inds.end(), t.begin(), t.end());
    d.vals.push_back(v);
  }
  return true;
}

bool read_file(const char *path, std::string &text) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return false;
  std::fseek(fp, 0, SEEK_END);
  const long sz = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(fp);
    return false;
  }
  text.resize((size_t)sz);
  const size_t got = sz ? std::fread(&text[0], 1, (size_t)sz, fp) : 0;
  std::fclose(fp);
  return got == (size_t)sz;
}

bool parse_u32(const char *b, const char *e, u32 &v) {
  const auto r = std::from_chars(b, e, v);
  return r.ec == std::errc() && r.ptr == e;
}

bool parse_value(const char *b, const char *e, double &v) {
  if (b == e) return false;
  // digits, a sign '-', or '.' start a number the stream reads the same way
  if (!((*b >= '0' && *b <= '9') || *b == '-' || *b == '.')) return false;
  if (*b == '-' && (e - b < 2 || !((b[1] >= '0' && b[1] <= '9') || b[1] == '.'))) return false;
  const auto r = std::from_chars(b, e, v, std::chars_format::general);
  return r.ec == std::errc() && r.ptr == e;
}

int worker_count(size_t bytes) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t by_size = bytes / (1u << 20) + 1;  // at least 1 MiB per chunk
  return (int)std::min<size_t>(std::min<unsigned>(hw, 16u), by_size);
}

template <class F>
void parallel_for(int n, F &&f) {
  if (n == 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(n);
  for (int k = 0; k < n; ++k) th.emplace_back([&f, k] { f(k); });
  for (auto &t : th) t.join();
}

// Returns false when the text needs the serial reader.
bool parse_parallel(const std::string &text, TnsData &d) {
  const char *p = text.data(), *end = p + text.size();
  auto skip = [&](const char *q) {
    while (q < end && is_ws(*q)) ++q;
    return q;
  };
  auto token_end = [&](const char *q) {
    while (q < end && !is_ws(*q)) ++q;
    return q;
  };
  // header
  const char *h0 = skip(p), *h1 = token_end(h0);
  const char *h2 = skip(h1), *h3 = token_end(h2);
  if (!parse_u32(h0, h1, d.order) || !parse_u32(h2, h3, d.dim)) return false;
  d.inds.clear();
  d.vals.clear();
  if (d.order == 0 || d.order > 64) return true;
  const u64 per = d.order + 1;
  const char *body = h3;
  const size_t blen = (size_t)(end - body);
  const int nw = worker_count(blen);
  // chunk starts, moved forward to a whitespace character
  std::vector<const char *> cs(nw + 1);
  cs[0] = body;
  cs[nw] = end;
  for (int k = 1; k < nw; ++k) {
    const char *q = body + blen * k / nw;
    q = std::max(q, cs[k - 1]);
    while (q < end && !is_ws(*q)) ++q;
    cs[k] = q;
  }
  std::vector<u64> cnt(nw + 1, 0);
  parallel_for(nw, [&](int k) {
    u64 c = 0;
    const char *q = cs[k];
    while (true) {
      while (q < cs[k + 1] && is_ws(*q)) ++q;
      if (q >= cs[k + 1]) break;
      while (q < cs[k + 1] && !is_ws(*q)) ++q;
      ++c;
    }
    cnt[k + 1] = c;
  });
  for (int k = 0; k < nw; ++k) cnt[k + 1] += cnt[k];
  const u64 K = cnt[nw];
  u64 ntup = K / per;
  // a complete last tuple whose value ends the file is dropped
  if (ntup > 0 && K % per == 0 && end > body && !is_ws(end[-1])) --ntup;
  d.inds.resize(ntup * d.order);
  d.vals.resize(ntup);
  std::vector<char> ok(nw, 1);
  parallel_for(nw, [&](int k) {
    u64 tok = cnt[k];
    const char *q = cs[k];
    while (tok < ntup * per) {
      while (q < cs[k + 1] && is_ws(*q)) ++q;
      if (q >= cs[k + 1]) break;
      const char *t0 = q;
      while (q < cs[k + 1] && !is_ws(*q)) ++q;
      const u64 i = tok / per, f = tok % per;
      if (f < d.order) {
        if (!parse_u32(t0, q, d.inds[i * d.order + f])) {
          ok[k] = 0;
          return;
        }
      } else if (!parse_value(t0, q, d.vals[i])) {
        ok[k] = 0;
        return;
      }
      ++tok;
    }
  });
  for (char c : ok)
    if (!c) return false;
  parallel_for(nw, [&](int k) {
    const u64 a = ntup * k / nw, b = ntup * (k + 1) / nw;
    for (u64 i = a; i < b; ++i)
      std::sort(d.inds.begin() + i * d.order, d.inds.begin() + (i + 1) * d.order);
  });
  return true;
}

bool read_tns(const char *path, TnsData &d) {
  std::string text;
  if (!read_file(path, text)) return false;
  if (parse_parallel(text, d)) return true;
  return read_serial(path, d);
}

struct SourceStamp {
  u64 size = 0;
  i64 mtime_ns = 0;
};

bool stamp_of(const char *path, SourceStamp &s) {
  struct stat st;
  if (::stat(path, &st) != 0) return false;
  s.size = (u64)st.st_size;
  s.mtime_ns = (i64)st.st_mtim.tv_sec * 1000000000ll + st.st_mtim.tv_nsec;
  return true;
}

// The tuples of the last file read with null output buffers, kept for the
// second call of the two-call protocol (same path, same stamp).
std::mutex g_last_mutex;
std::string g_last_path;
SourceStamp g_last_stamp;
TnsData g_last;

bool take_last(const char *path, const SourceStamp &s, TnsData &d) {
  std::lock_guard<std::mutex> lk(g_last_mutex);
  if (g_last_path != path || g_last_stamp.size != s.size || g_last_stamp.mtime_ns != s.mtime_ns)
    return false;
  d = std::move(g_last);
  g_last = TnsData{};
  g_last_path.clear();
  return true;
}

void keep_last(const char *path, const SourceStamp &s, TnsData &&d) {
  std::lock_guard<std::mutex> lk(g_last_mutex);
  g_last_path = path;
  g_last_stamp = s;
  g_last = std::move(d);
}

constexpr char kMagic[8] = {'S', 'P', 'T', 'N', 'S', 'B', 'I', 'N'};
constexpr u32 kVersion = 1;

struct CacheHeader {
  char magic[8];
  u32 version, order, dim, pad;
  u64 nnz, src_size;
  i64 src_mtime_ns;
};

bool read_cache(const char *cache, const SourceStamp &s, TnsData &d) {
  FILE *fp = std::fopen(cache, "rb");
  if (!fp) return false;
  CacheHeader h;
  bool good = std::fread(&h, sizeof h, 1, fp) == 1 &&
              std::memcmp(h.magic, kMagic, 8) == 0 && h.version == kVersion &&
              h.src_size == s.size && h.src_mtime_ns == s.mtime_ns && h.order <= 64;
  if (good) {
    d.order = h.order;
    d.dim = h.dim;
    d.inds.resize(h.nnz * h.order);
    d.vals.resize(h.nnz);
    good = std::fread(d.inds.data(), sizeof(u32), d.inds.size(), fp) == d.inds.size() &&
           std::fread(d.vals.data(), sizeof(double), d.vals.size(), fp) == d.vals.size();
  }
  std::fclose(fp);
  return good;
}

bool write_cache(const char *cache, const SourceStamp &s, const TnsData &d) {
  const std::string tmp = std::string(cache) + ".tmp";
  FILE *fp = std::fopen(tmp.c_str(), "wb");
  if (!fp) return false;
  CacheHeader h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.order = d.order;
  h.dim = d.dim;
  h.nnz = d.vals.size();
  h.src_size = s.size;
  h.src_mtime_ns = s.mtime_ns;
  bool good = std::fwrite(&h, sizeof h, 1, fp) == 1 &&
              std::fwrite(d.inds.data(), sizeof(u32), d.inds.size(), fp) == d.inds.size() &&
              std::fwrite(d.vals.data(), sizeof(double), d.vals.size(), fp) == d.vals.size();
  good = (std::fclose(fp) == 0) && good;
  if (!good || std::rename(tmp.c_str(), cache) != 0) {
    std::remove(tmp.c_str());
    return false;
  }
  return true;
}

int deliver(TnsData &d, const char *path, const SourceStamp &s, uint32_t *order, uint32_t *dim,
            uint64_t *nnz, uint32_t *inds, double *vals) {
  *order = d.order;
  *dim = d.dim;
  *nnz = d.vals.size();
  if (!inds && !vals) {
    keep_last(path, s, std::move(d));
    return SP_OK;
  }
  if (inds) std::copy(d.inds.begin(), d.inds.end(), inds);
  if (vals) std::copy(d.vals.begin(), d.vals.end(), vals);
  return SP_OK;
}

int fail(int code, const std::string &msg) {
  set_last_error(msg);
  return code;
}

int load(const char *path, const char *cache, uint32_t *order, uint32_t *dim, uint64_t *nnz,
         uint32_t *inds, double *vals) {
  if (!path || !order || !dim || !nnz) return fail(SP_ERR_ARG, "null argument");
  try {
    SourceStamp s;
    if (!stamp_of(path, s)) return fail(SP_ERR_IO, std::string("cannot open ") + path);
    TnsData d;
    if (take_last(path, s, d))
      return deliver(d, path, s, order, dim, nnz, inds, vals);
    if (cache && read_cache(cache, s, d))
      return deliver(d, path, s, order, dim, nnz, inds, vals);
    if (!read_tns(path, d)) return fail(SP_ERR_IO, std::string("cannot read ") + path);
    if (d.order == 0 || d.order > 64)
      return fail(SP_ERR_IO, std::string("bad header (order) in ") + path);
    if (cache && !write_cache(cache, s, d))
      return fail(SP_ERR_IO, std::string("cannot write the tuple cache ") + cache);
    return deliver(d, path, s, order, dim, nnz, inds, vals);
  } catch (const std::bad_alloc &) {
    return fail(SP_ERR_OOM, "host allocation failed");
  }
}

}  // namespace

}  // namespace spamd

extern "C" {

// src/utils/loader.cpp:8-36
int sp_load_tns(const char *path, uint32_t *order, uint32_t *dim, uint64_t *nnz,
                uint32_t *inds, double *vals) {
  return spamd::load(path, nullptr, order, dim, nnz, inds, vals);
}

int sp_load_tns_cached(const char *path, const char *cache, uint32_t *order, uint32_t *dim,
                       uint64_t *nnz, uint32_t *inds, double *vals) {
  return spamd::load(path, cache, order, dim, nnz, inds, vals);
}

}  // extern "C"
