// textio.hip - the reference's GNU checkpoint file of a field, written and
// read on the GPU (sps_fd_checkpoint_text, sps_fd_checkpoint_read).
//
// Format (io/src/writegp.cpp: print_gp_header :427-465, save_grid :780-803;
// io/src/readgp.cpp: read_header, read_block): a header line
//   "dim n0 n1 [n2] left0 right0 ... index @\n"
// then every value as "%.10E " in x-fastest order, "\n" after each x row.
//
// Writing: pass 1 formats every value in registers (decimal.hpp) and keeps
// only its length; the few values the exact fast formatter declines (near a
// rounding tie, subnormal, inf, nan; about 2e-4 of random field values) are
// formatted on the host with the C library and their lengths patched in.
// Per-block length sums are scanned into byte offsets (plus one newline per
// completed row) and pass 2 formats again and writes each value's bytes at
// its offset.  The field is processed in row-aligned chunks so the text
// buffer on the device stays bounded.
//
// Reading: token starts (non-space after space) are counted per 4 KiB block,
// scanned into token indices, and every token is parsed by the exact fast
// path (decimal.hpp); tokens it declines are parsed on the host by strtod,
// as fscanf("%lf") would.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "decimal.hpp"
#include "plan.hpp"
#include "textio.hpp"

namespace sps {
namespace {

constexpr int kThreads = 256;
constexpr int kValsPerThread = 4;
constexpr int kValsPerBlock = kThreads * kValsPerThread;  // writer
constexpr int kBytesPerThread = 16;
constexpr int kBytesPerBlock = kThreads * kBytesPerThread;  // reader
constexpr int kMaxText = 20;  // "-1.2345678901E+100 " is 19 bytes
constexpr int64_t kChunkValues = int64_t(1) << 24;

// exclusive prefix sum of v over the 256 threads of a block; *total gets the
// block's sum
template <typename U>
__device__ U block_exclusive_scan(U v, U* total) {
  __shared__ U sh[kThreads];
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < kThreads; d <<= 1) {
    U add = t >= d ? sh[t - d] : U(0);
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  const U incl = sh[t];
  *total = sh[kThreads - 1];
  __syncthreads();
  return incl - v;
}

// How values are laid out in a text file: "%.<P>E" numbers (P a template
// argument), optionally with the blank flag, each followed by sep (after the
// last of a row too when sep_last), one row of X values per line.  Rows can
// carry a prefix (position columns, slice headers): extra[r] is the number of
// non-value bytes before the values of chunk row r (prefixes of rows <= r plus
// the newlines of rows < r; r itself when extra is NULL), and row r's prefix
// is pre_str[pre_off[r] .. pre_off[r+1]).
struct TextLayout {
  int space;
  char sep;
  int sep_last;
  const uint64_t* extra;
  const uint32_t* pre_off;
  const char* pre_str;
};

__device__ __forceinline__ bool has_sep(const TextLayout& L, int64_t i,
                                        int64_t X) {
  return L.sep_last || (i + 1) % X != 0;
}

template <typename T, int P>
__global__ __launch_bounds__(kThreads) void fmt_len(
    const T* __restrict__ f, int64_t n, int64_t X, TextLayout L,
    uint8_t* __restrict__ len, uint32_t* __restrict__ flag_idx,
    unsigned* __restrict__ nflag) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  char buf[24];
  const int k = dec::format_e<P>((double)f[i], buf, L.space != 0);
  if (k == 0) {
    const unsigned slot = atomicAdd(nflag, 1u);  // capacity n
    flag_idx[slot] = (uint32_t)i;
    len[i] = 0;
  } else {
    len[i] = (uint8_t)(k + (has_sep(L, i, X) ? 1 : 0));
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void gather_values(
    const T* __restrict__ f, const uint32_t* __restrict__ idx, unsigned n,
    double* __restrict__ out) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) out[i] = (double)f[idx[i]];
}

__global__ __launch_bounds__(kThreads) void patch_len(
    uint8_t* __restrict__ len, const uint32_t* __restrict__ idx,
    const uint8_t* __restrict__ l, unsigned n) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) len[idx[i]] = l[i];
}

__global__ __launch_bounds__(kThreads) void len_block_sum(
    const uint8_t* __restrict__ len, int64_t n, uint32_t* __restrict__ bsum) {
  const int64_t i0 = (int64_t)blockIdx.x * kValsPerBlock +
                     (int64_t)threadIdx.x * kValsPerThread;
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < kValsPerThread; ++q)
    if (i0 + q < n) s += len[i0 + q];
  uint32_t total;
  block_exclusive_scan<uint32_t>(s, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// boff[b]: bytes of the values before value b*kValsPerBlock in the chunk's
// text (the layout's non-value bytes not included).  Flagged values (sorted
// fidx) take their text, separator included, from fstr.
template <typename T, int P>
__global__ __launch_bounds__(kThreads) void fmt_write(
    const T* __restrict__ f, int64_t n, int64_t X, TextLayout L,
    const uint8_t* __restrict__ len, const uint64_t* __restrict__ boff,
    const uint32_t* __restrict__ fidx, const uint32_t* __restrict__ foff,
    const char* __restrict__ fstr, unsigned nf, char* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * kValsPerBlock;
  const int64_t i0 = base + (int64_t)threadIdx.x * kValsPerThread;
  uint32_t l[kValsPerThread];
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < kValsPerThread; ++q) {
    l[q] = i0 + q < n ? len[i0 + q] : 0;
    s += l[q];
  }
  uint32_t total;
  const uint32_t excl = block_exclusive_scan<uint32_t>(s, &total);
  uint64_t off = boff[blockIdx.x] + excl;
#pragma unroll
  for (int q = 0; q < kValsPerThread; ++q) {
    const int64_t i = i0 + q;
    if (i >= n) break;
    const int64_t r = i / X;
    char* o = out + off + (L.extra ? L.extra[r] : (uint64_t)r);
    if (L.pre_off && i % X == 0) {
      const uint32_t a = L.pre_off[r], b = L.pre_off[r + 1];
      for (uint32_t c = a; c < b; ++c) o[(int64_t)c - (int64_t)b] = L.pre_str[c];
    }
    char buf[24];
    int k = dec::format_e<P>((double)f[i], buf, L.space != 0);
    if (k > 0) {
      for (int c = 0; c < k; ++c) o[c] = buf[c];
      if (has_sep(L, i, X)) o[k] = L.sep;
    } else {
      // binary search of the sorted flagged indices
      unsigned lo = 0, hi = nf;
      while (lo + 1 < hi) {
        const unsigned mid = (lo + hi) / 2;
        if (fidx[mid] <= (uint32_t)i) lo = mid; else hi = mid;
      }
      const uint32_t a = foff[lo], b = foff[lo + 1];
      for (uint32_t c = a; c < b; ++c) o[c - a] = fstr[c];
    }
    if ((i + 1) % X == 0) o[l[q]] = '\n';
    off += l[q];
  }
}

// XDR doubles (RFC 4506: big-endian IEEE 754) from the field, two per thread.
template <typename T>
__global__ __launch_bounds__(kThreads) void xdr_pack(const T* __restrict__ f,
                                                     int64_t n,
                                                     uint64_t* __restrict__ out) {
  const int64_t i = ((int64_t)blockIdx.x * kThreads + threadIdx.x) * 2;
  if (i + 1 < n) {
    const double a = (double)f[i], b = (double)f[i + 1];
    out[i] = __builtin_bswap64((uint64_t)__double_as_longlong(a));
    out[i + 1] = __builtin_bswap64((uint64_t)__double_as_longlong(b));
  } else if (i < n) {
    out[i] = __builtin_bswap64((uint64_t)__double_as_longlong((double)f[i]));
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void xdr_unpack(
    const uint64_t* __restrict__ in, int64_t n, T* __restrict__ f) {
  const int64_t i = ((int64_t)blockIdx.x * kThreads + threadIdx.x) * 2;
  if (i + 1 < n) {
    f[i] = (T)__longlong_as_double((long long)__builtin_bswap64(in[i]));
    f[i + 1] = (T)__longlong_as_double((long long)__builtin_bswap64(in[i + 1]));
  } else if (i < n) {
    f[i] = (T)__longlong_as_double((long long)__builtin_bswap64(in[i]));
  }
}

__global__ __launch_bounds__(kThreads) void tok_count(
    const char* __restrict__ s, int64_t m, uint32_t* __restrict__ bcount) {
  const int64_t b0 = (int64_t)blockIdx.x * kBytesPerBlock +
                     (int64_t)threadIdx.x * kBytesPerThread;
  uint32_t c = 0;
  for (int q = 0; q < kBytesPerThread; ++q) {
    const int64_t i = b0 + q;
    if (i >= m) break;
    if (!dec::is_space(s[i]) && (i == 0 || dec::is_space(s[i - 1]))) ++c;
  }
  uint32_t total;
  block_exclusive_scan<uint32_t>(c, &total);
  if (threadIdx.x == 0) bcount[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void tok_parse(
    const char* __restrict__ s, int64_t m, const uint64_t* __restrict__ boff,
    int64_t nvals, T* __restrict__ field, uint32_t* __restrict__ flag_tok,
    uint64_t* __restrict__ flag_pos, unsigned* __restrict__ nflag) {
  const int64_t b0 = (int64_t)blockIdx.x * kBytesPerBlock +
                     (int64_t)threadIdx.x * kBytesPerThread;
  uint32_t c = 0;
  for (int q = 0; q < kBytesPerThread; ++q) {
    const int64_t i = b0 + q;
    if (i >= m) break;
    if (!dec::is_space(s[i]) && (i == 0 || dec::is_space(s[i - 1]))) ++c;
  }
  uint32_t total;
  uint64_t tok = boff[blockIdx.x] + block_exclusive_scan<uint32_t>(c, &total);
  for (int q = 0; q < kBytesPerThread; ++q) {
    const int64_t i = b0 + q;
    if (i >= m) break;
    if (dec::is_space(s[i]) || (i > 0 && !dec::is_space(s[i - 1]))) continue;
    if ((int64_t)tok < nvals) {
      int len = 0;
      while (i + len < m && len < 64 && !dec::is_space(s[i + len])) ++len;
      double v;
      if (len < 64 && dec::parse_fast(s + i, len, &v)) {
        field[tok] = (T)v;
      } else {
        const unsigned slot = atomicAdd(nflag, 1u);  // capacity nvals
        flag_tok[slot] = (uint32_t)tok;
        flag_pos[slot] = (uint64_t)i;
      }
    }
    ++tok;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void scatter_values(
    T* __restrict__ field, const uint32_t* __restrict__ idx,
    const double* __restrict__ v, unsigned n) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) field[idx[i]] = (T)v[i];
}

// device buffers of one call, freed on every exit path
struct DevBufs {
  std::vector<void*> ptrs;
  ~DevBufs() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename P>
  int alloc(P** p, size_t bytes) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
    if (e != hipSuccess) {
      set_error("hipMalloc of %zu bytes failed: %s", bytes,
                hipGetErrorString(e));
      return SPS_E_NOMEM;
    }
    ptrs.push_back(q);
    *p = (P*)q;
    return SPS_OK;
  }
};

std::string gp_header(const sps_fd_desc& d, int index, const double* origin) {
  // print_gp_header (io/src/writegp.cpp:427-465): "%d " dimension,
  // "%d " counts, "%lf %lf " domain bounds, "%d" index, " %c" '@', "\n"
  std::string h;
  char buf[128];
  std::snprintf(buf, sizeof(buf), "%d ", d.dim);
  h += buf;
  for (int a = 0; a < d.dim; ++a) {
    std::snprintf(buf, sizeof(buf), "%d ", d.n[a]);
    h += buf;
  }
  for (int a = 0; a < d.dim; ++a) {
    const double left = origin ? origin[a] : 0.0;
    std::snprintf(buf, sizeof(buf), "%lf %lf ", left,
                  left + (d.n[a] - 1) * d.h);
    h += buf;
  }
  std::snprintf(buf, sizeof(buf), "%d %c\n", index, '@');
  h += buf;
  return h;
}

unsigned nblocks(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

// A text file layout on the host side: the number format, the separators,
// the file header and tail, and the prefix of each row (global row index g,
// appended to out), if any.
struct TextFormat {
  int prec = 10;  // 10 ("%.10E") or 6 ("% 10lE")
  bool space = false;
  char sep = ' ';
  bool sep_last = true;
  std::string header, tail;
  std::function<void(int64_t g, std::string& out)> prefix;
};

template <typename T, int P>
void launch_fmt_len(const T* f, int64_t cnt, int64_t X, const TextLayout& L,
                    uint8_t* len, uint32_t* fidx, unsigned* nflag,
                    hipStream_t st) {
  fmt_len<T, P><<<nblocks(cnt, kThreads), kThreads, 0, st>>>(f, cnt, X, L, len,
                                                             fidx, nflag);
}

template <typename T, int P>
void launch_fmt_write(const T* f, int64_t cnt, int64_t X, const TextLayout& L,
                      const uint8_t* len, const uint64_t* boff,
                      const uint32_t* fidx, const uint32_t* foff,
                      const char* fstr, unsigned nf, char* out,
                      hipStream_t st) {
  fmt_write<T, P><<<nblocks(cnt, kValsPerBlock), kThreads, 0, st>>>(
      f, cnt, X, L, len, boff, fidx, foff, fstr, nf, out);
}

template <typename T>
int write_text(const FieldView& v, const TextFormat& fm, char* text,
               size_t capacity, size_t* length) {
  hipStream_t st = v.stream;
  const T* field = (const T*)v.data;
  const int64_t X = v.d.n[0];
  const int64_t nsites = (int64_t)v.nsites;
  const int64_t rows_per_chunk = std::max<int64_t>(1, kChunkValues / X);
  const int64_t chunk_cap = std::min<int64_t>(nsites, rows_per_chunk * X);
  const unsigned nb_cap = nblocks(chunk_cap, kValsPerBlock);
  const bool p6 = fm.prec == 6;

  DevBufs db;
  uint8_t* d_len;
  uint32_t *d_fidx, *d_bsum;
  unsigned* d_nflag;
  uint64_t* d_boff;
  SPS_TRY(db.alloc(&d_len, chunk_cap));
  SPS_TRY(db.alloc(&d_fidx, chunk_cap * sizeof(uint32_t)));
  SPS_TRY(db.alloc(&d_nflag, sizeof(unsigned)));
  SPS_TRY(db.alloc(&d_bsum, nb_cap * sizeof(uint32_t)));
  SPS_TRY(db.alloc(&d_boff, nb_cap * sizeof(uint64_t)));
  char* d_out = nullptr;
  const size_t value_room = (size_t)chunk_cap * kMaxText;

  size_t total = fm.header.size();
  if (text) {
    if (capacity < total) {
      set_error("text file: capacity %zu is below the header alone", capacity);
      return SPS_E_INVALID;
    }
    std::memcpy(text, fm.header.data(), fm.header.size());
  }
  std::vector<uint32_t> hidx;
  std::vector<double> hval;
  std::vector<uint32_t> hbsum(nb_cap);
  std::vector<uint64_t> hboff(nb_cap);
  std::string pre;
  std::vector<uint32_t> pre_off;
  std::vector<uint64_t> extra;
  size_t out_cap = 0;
  for (int64_t start = 0; start < nsites; start += chunk_cap) {
    const int64_t cnt = std::min<int64_t>(chunk_cap, nsites - start);
    const int64_t rows = cnt / X;
    const T* f = field + start;
    const unsigned nb = nblocks(cnt, kValsPerBlock);
    DevBufs chunkbufs;
    TextLayout L{fm.space ? 1 : 0, fm.sep, fm.sep_last ? 1 : 0, nullptr, nullptr,
                 nullptr};
    uint64_t extra_total = (uint64_t)rows;  // the newlines
    if (fm.prefix) {
      pre.clear();
      pre_off.assign((size_t)rows + 1, 0);
      extra.assign((size_t)rows, 0);
      uint64_t acc = 0;
      for (int64_t r = 0; r < rows; ++r) {
        pre_off[r] = (uint32_t)pre.size();
        fm.prefix(start / X + r, pre);
        const uint64_t pl = pre.size() - pre_off[r];
        extra[r] = acc + pl;
        acc += pl + 1;
      }
      pre_off[rows] = (uint32_t)pre.size();
      extra_total = acc;
      uint64_t* d_extra;
      uint32_t* d_preoff;
      char* d_pre;
      SPS_TRY(chunkbufs.alloc(&d_extra, rows * sizeof(uint64_t)));
      SPS_TRY(chunkbufs.alloc(&d_preoff, (rows + 1) * sizeof(uint32_t)));
      SPS_TRY(chunkbufs.alloc(&d_pre, pre.size()));
      SPS_HIP_TRY(hipMemcpyAsync(d_extra, extra.data(), rows * sizeof(uint64_t),
                                 hipMemcpyHostToDevice, st));
      SPS_HIP_TRY(hipMemcpyAsync(d_preoff, pre_off.data(),
                                 (rows + 1) * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, st));
      if (!pre.empty())
        SPS_HIP_TRY(hipMemcpyAsync(d_pre, pre.data(), pre.size(),
                                   hipMemcpyHostToDevice, st));
      L.extra = d_extra;
      L.pre_off = d_preoff;
      L.pre_str = d_pre;
    }
    SPS_HIP_TRY(hipMemsetAsync(d_nflag, 0, sizeof(unsigned), st));
    if (p6)
      launch_fmt_len<T, 6>(f, cnt, X, L, d_len, d_fidx, d_nflag, st);
    else
      launch_fmt_len<T, 10>(f, cnt, X, L, d_len, d_fidx, d_nflag, st);
    unsigned nf = 0;
    SPS_HIP_TRY(hipMemcpyAsync(&nf, d_nflag, sizeof(unsigned),
                               hipMemcpyDeviceToHost, st));
    SPS_HIP_TRY(hipStreamSynchronize(st));
    // the declined values, formatted by the C library
    uint32_t *d_sidx = nullptr, *d_soff = nullptr;
    char* d_sstr = nullptr;
    DevBufs flagbufs;
    if (nf > 0) {
      hidx.resize(nf);
      hval.resize(nf);
      double* d_val;
      SPS_TRY(flagbufs.alloc(&d_val, nf * sizeof(double)));
      SPS_HIP_TRY(hipMemcpyAsync(hidx.data(), d_fidx, nf * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, st));
      gather_values<T><<<nblocks(nf, kThreads), kThreads, 0, st>>>(f, d_fidx, nf,
                                                                   d_val);
      SPS_HIP_TRY(hipMemcpyAsync(hval.data(), d_val, nf * sizeof(double),
                                 hipMemcpyDeviceToHost, st));
      SPS_HIP_TRY(hipStreamSynchronize(st));
      std::vector<std::pair<uint32_t, double>> pv(nf);
      for (unsigned k = 0; k < nf; ++k) pv[k] = {hidx[k], hval[k]};
      std::sort(pv.begin(), pv.end());
      std::vector<uint32_t> sidx(nf), soff(nf + 1);
      std::vector<uint8_t> slen(nf);
      std::string sstr;
      char buf[64];
      for (unsigned k = 0; k < nf; ++k) {
        int w = p6 ? std::snprintf(buf, sizeof(buf), "% 10lE", pv[k].second)
                   : std::snprintf(buf, sizeof(buf), "%.10E", pv[k].second);
        if (fm.sep_last || (pv[k].first + 1) % X != 0) buf[w++] = fm.sep;
        sidx[k] = pv[k].first;
        soff[k] = (uint32_t)sstr.size();
        slen[k] = (uint8_t)w;
        sstr.append(buf, (size_t)w);
      }
      soff[nf] = (uint32_t)sstr.size();
      uint8_t* d_slen;
      SPS_TRY(flagbufs.alloc(&d_sidx, nf * sizeof(uint32_t)));
      SPS_TRY(flagbufs.alloc(&d_soff, (nf + 1) * sizeof(uint32_t)));
      SPS_TRY(flagbufs.alloc(&d_slen, nf));
      SPS_TRY(flagbufs.alloc(&d_sstr, sstr.size()));
      SPS_HIP_TRY(hipMemcpyAsync(d_sidx, sidx.data(), nf * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, st));
      SPS_HIP_TRY(hipMemcpyAsync(d_soff, soff.data(), (nf + 1) * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, st));
      SPS_HIP_TRY(hipMemcpyAsync(d_slen, slen.data(), nf, hipMemcpyHostToDevice,
                                 st));
      SPS_HIP_TRY(hipMemcpyAsync(d_sstr, sstr.data(), sstr.size(),
                                 hipMemcpyHostToDevice, st));
      patch_len<<<nblocks(nf, kThreads), kThreads, 0, st>>>(d_len, d_sidx, d_slen,
                                                            nf);
      SPS_HIP_TRY(hipStreamSynchronize(st));  // host vectors above go away
    }
    len_block_sum<<<nb, kThreads, 0, st>>>(d_len, cnt, d_bsum);
    SPS_HIP_TRY(hipMemcpyAsync(hbsum.data(), d_bsum, nb * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, st));
    SPS_HIP_TRY(hipStreamSynchronize(st));
    uint64_t acc = 0;
    for (unsigned b = 0; b < nb; ++b) {
      hboff[b] = acc;
      acc += hbsum[b];
    }
    const uint64_t chunk_bytes = acc + extra_total;
    if (text) {
      if (total + chunk_bytes > capacity) {
        set_error("text file: capacity %zu too small", capacity);
        return SPS_E_INVALID;
      }
      const size_t need = std::max<size_t>(value_room + (size_t)extra_total,
                                           (size_t)chunk_bytes) + 16;
      if (need > out_cap) {
        if (d_out) {
          (void)hipFree(d_out);
          db.ptrs.erase(std::find(db.ptrs.begin(), db.ptrs.end(), (void*)d_out));
          d_out = nullptr;
        }
        SPS_TRY(db.alloc(&d_out, need));
        out_cap = need;
      }
      SPS_HIP_TRY(hipMemcpyAsync(d_boff, hboff.data(), nb * sizeof(uint64_t),
                                 hipMemcpyHostToDevice, st));
      if (p6)
        launch_fmt_write<T, 6>(f, cnt, X, L, d_len, d_boff, d_sidx, d_soff,
                               d_sstr, nf, d_out, st);
      else
        launch_fmt_write<T, 10>(f, cnt, X, L, d_len, d_boff, d_sidx, d_soff,
                                d_sstr, nf, d_out, st);
      SPS_HIP_TRY(hipMemcpyAsync(text + total, d_out, chunk_bytes,
                                 hipMemcpyDeviceToHost, st));
      SPS_HIP_TRY(hipStreamSynchronize(st));
    }
    total += chunk_bytes;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("text file kernels failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  if (text) {
    if (total + fm.tail.size() > capacity) {
      set_error("text file: capacity %zu too small", capacity);
      return SPS_E_INVALID;
    }
    std::memcpy(text + total, fm.tail.data(), fm.tail.size());
  }
  total += fm.tail.size();
  if (text && total < capacity) text[total] = 0;
  *length = total;
  return SPS_OK;
}

// The plot-file layouts (io/src/writegp.cpp save_grid_plotting :470-507,
// io/src/writecsv.cpp save_grid_plotting): "% 10lE" values; GNU matrix rows
// start with the y position "% 13.2f " and every z slice with a line of its
// index "% 13d " and the x positions, slices separated by a blank line, and
// the file ends with an empty line.
TextFormat plot_format(const sps_fd_desc& d, int format, const double* origin) {
  TextFormat fm;
  fm.prec = 6;
  fm.space = true;
  if (format == SPS_PLOT_CSV) {
    fm.sep = ',';
    fm.sep_last = false;
    return fm;
  }
  fm.sep = ' ';
  fm.sep_last = true;
  fm.tail = "\n";
  const double x0 = origin ? origin[0] : 0.0, y0 = origin ? origin[1] : 0.0;
  const double h = d.h;
  const int X = d.n[0], Y = d.n[1];
  fm.prefix = [x0, y0, h, X, Y](int64_t g, std::string& out) {
    char buf[64];
    const int64_t j = g % Y, k = g / Y;
    if (j == 0) {
      if (k > 0) out += '\n';
      std::snprintf(buf, sizeof(buf), "% 13d ", (int)k);
      out += buf;
      for (int i = 0; i < X; ++i) {
        std::snprintf(buf, sizeof(buf), "% 13.2f ", x0 + i * h);
        out += buf;
      }
      out += '\n';
    }
    std::snprintf(buf, sizeof(buf), "% 13.2f ", y0 + (double)j * h);
    out += buf;
  };
  return fm;
}

// Header as read_header (io/src/readgp.cpp) reads it; returns the bytes
// consumed, or -1.
long parse_header(const sps_fd_desc& d, const char* text, size_t length,
                  int* index) {
  size_t hl = 0;
  while (hl < length && hl < 4096 && text[hl] != '@') ++hl;
  if (hl >= length || text[hl] != '@') {
    set_error("checkpoint header: no '@' (subdomain headers are not "
              "supported) within the first 4096 bytes");
    return -1;
  }
  std::string h(text, hl);
  const char* p = h.c_str();
  int used = 0, dim = 0;
  if (std::sscanf(p, "%d%n", &dim, &used) != 1 || dim != d.dim) {
    set_error("checkpoint header: dimension does not match the plan");
    return -1;
  }
  p += used;
  for (int a = 0; a < dim; ++a) {
    int c = 0;
    if (std::sscanf(p, "%d%n", &c, &used) != 1 || c != d.n[a]) {
      set_error("checkpoint header: count of axis %d does not match the plan",
                a);
      return -1;
    }
    p += used;
  }
  for (int a = 0; a < dim; ++a) {
    double l, r;
    if (std::sscanf(p, "%lf %lf%n", &l, &r, &used) != 2) {
      set_error("checkpoint header: interval of axis %d unreadable", a);
      return -1;
    }
    p += used;
  }
  int idx = 0;
  if (std::sscanf(p, "%d%n", &idx, &used) != 1) {
    set_error("checkpoint header: index unreadable");
    return -1;
  }
  p += used;
  while (*p == ' ' || *p == '\t') ++p;
  if (*p != 0) {
    set_error("checkpoint header: unexpected text before '@'");
    return -1;
  }
  if (index) *index = idx;
  return (long)hl + 1;
}

template <typename T>
int read_text(const FieldView& v, const char* body, int64_t m) {
  hipStream_t st = v.stream;
  const int64_t nvals = (int64_t)v.nsites;
  const unsigned nb = nblocks(std::max<int64_t>(m, 1), kBytesPerBlock);
  DevBufs db;
  char* d_s;
  uint32_t *d_bcount, *d_ftok;
  uint64_t *d_boff, *d_fpos;
  unsigned* d_nflag;
  SPS_TRY(db.alloc(&d_s, (size_t)m));
  SPS_TRY(db.alloc(&d_bcount, nb * sizeof(uint32_t)));
  SPS_TRY(db.alloc(&d_boff, nb * sizeof(uint64_t)));
  SPS_TRY(db.alloc(&d_ftok, nvals * sizeof(uint32_t)));
  SPS_TRY(db.alloc(&d_fpos, nvals * sizeof(uint64_t)));
  SPS_TRY(db.alloc(&d_nflag, sizeof(unsigned)));
  if (m > 0)
    SPS_HIP_TRY(hipMemcpyAsync(d_s, body, (size_t)m, hipMemcpyHostToDevice, st));
  tok_count<<<nb, kThreads, 0, st>>>(d_s, m, d_bcount);
  std::vector<uint32_t> bc(nb);
  std::vector<uint64_t> bo(nb);
  SPS_HIP_TRY(hipMemcpyAsync(bc.data(), d_bcount, nb * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, st));
  SPS_HIP_TRY(hipStreamSynchronize(st));
  uint64_t acc = 0;
  for (unsigned b = 0; b < nb; ++b) {
    bo[b] = acc;
    acc += bc[b];
  }
  if ((int64_t)acc < nvals) {
    set_error("checkpoint: %llu values for a field of %lld sites",
              (unsigned long long)acc, (long long)nvals);
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipMemcpyAsync(d_boff, bo.data(), nb * sizeof(uint64_t),
                             hipMemcpyHostToDevice, st));
  SPS_HIP_TRY(hipMemsetAsync(d_nflag, 0, sizeof(unsigned), st));
  T* field = (T*)v.data;
  tok_parse<T><<<nb, kThreads, 0, st>>>(d_s, m, d_boff, nvals, field, d_ftok,
                                        d_fpos, d_nflag);
  unsigned nf = 0;
  SPS_HIP_TRY(hipMemcpyAsync(&nf, d_nflag, sizeof(unsigned),
                             hipMemcpyDeviceToHost, st));
  SPS_HIP_TRY(hipStreamSynchronize(st));
  if (nf > 0) {
    // tokens the fast path declined: strtod, as fscanf("%lf") parses them
    std::vector<uint32_t> tok(nf);
    std::vector<uint64_t> pos(nf);
    std::vector<double> val(nf);
    SPS_HIP_TRY(hipMemcpyAsync(tok.data(), d_ftok, nf * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, st));
    SPS_HIP_TRY(hipMemcpyAsync(pos.data(), d_fpos, nf * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, st));
    SPS_HIP_TRY(hipStreamSynchronize(st));
    for (unsigned k = 0; k < nf; ++k) {
      int64_t e = (int64_t)pos[k];
      while (e < m && !dec::is_space(body[e])) ++e;
      std::string t(body + pos[k], (size_t)(e - (int64_t)pos[k]));
      char* end = nullptr;
      val[k] = std::strtod(t.c_str(), &end);
      if (end == t.c_str() || *end != 0) {
        set_error("checkpoint: value %u (\"%.40s\") is not a number", tok[k],
                  t.c_str());
        return SPS_E_INVALID;
      }
    }
    uint32_t* d_idx;
    double* d_val;
    DevBufs fb;
    SPS_TRY(fb.alloc(&d_idx, nf * sizeof(uint32_t)));
    SPS_TRY(fb.alloc(&d_val, nf * sizeof(double)));
    SPS_HIP_TRY(hipMemcpyAsync(d_idx, tok.data(), nf * sizeof(uint32_t),
                               hipMemcpyHostToDevice, st));
    SPS_HIP_TRY(hipMemcpyAsync(d_val, val.data(), nf * sizeof(double),
                               hipMemcpyHostToDevice, st));
    scatter_values<T><<<nblocks(nf, kThreads), kThreads, 0, st>>>(field, d_idx,
                                                                  d_val, nf);
    SPS_HIP_TRY(hipStreamSynchronize(st));
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("checkpoint read kernels failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

// XDR file header (io/src/writexdr.cpp print_xdr_header :45-87): int dim,
// int counts, per axis the domain (left, right), int index, int dim, per axis
// the interval (left, right); ints as 4 big-endian bytes, doubles as 8.
std::string xdr_header(const sps_fd_desc& d, int index, const double* origin) {
  std::string h;
  auto put_int = [&h](int v) {
    const uint32_t u = (uint32_t)v;
    for (int b = 3; b >= 0; --b) h += (char)(u >> (8 * b));
  };
  auto put_double = [&h](double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    for (int b = 7; b >= 0; --b) h += (char)(u >> (8 * b));
  };
  auto bounds = [&](int a) {
    const double left = origin ? origin[a] : 0.0;
    put_double(left);
    put_double(left + (d.n[a] - 1) * d.h);
  };
  put_int(d.dim);
  for (int a = 0; a < d.dim; ++a) put_int(d.n[a]);
  for (int a = 0; a < d.dim; ++a) bounds(a);
  put_int(index);
  put_int(d.dim);
  for (int a = 0; a < d.dim; ++a) bounds(a);
  return h;
}

constexpr int64_t kXdrChunk = int64_t(1) << 25;  // values per staged copy

template <typename T>
int write_xdr(const FieldView& v, const std::string& header,
              unsigned char* data) {
  hipStream_t st = v.stream;
  const T* field = (const T*)v.data;
  const int64_t n = (int64_t)v.nsites;
  std::memcpy(data, header.data(), header.size());
  DevBufs db;
  uint64_t* d_buf;
  const int64_t chunk = std::min<int64_t>(n, kXdrChunk);
  SPS_TRY(db.alloc(&d_buf, (size_t)chunk * 8));
  for (int64_t s0 = 0; s0 < n; s0 += chunk) {
    const int64_t cnt = std::min<int64_t>(chunk, n - s0);
    xdr_pack<T><<<nblocks((cnt + 1) / 2, kThreads), kThreads, 0, st>>>(
        field + s0, cnt, d_buf);
    SPS_HIP_TRY(hipMemcpyAsync(data + header.size() + (size_t)s0 * 8, d_buf,
                               (size_t)cnt * 8, hipMemcpyDeviceToHost, st));
    SPS_HIP_TRY(hipStreamSynchronize(st));
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("XDR kernels failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template <typename T>
int read_xdr(const FieldView& v, const unsigned char* body) {
  hipStream_t st = v.stream;
  T* field = (T*)v.data;
  const int64_t n = (int64_t)v.nsites;
  DevBufs db;
  uint64_t* d_buf;
  const int64_t chunk = std::min<int64_t>(n, kXdrChunk);
  SPS_TRY(db.alloc(&d_buf, (size_t)chunk * 8));
  for (int64_t s0 = 0; s0 < n; s0 += chunk) {
    const int64_t cnt = std::min<int64_t>(chunk, n - s0);
    SPS_HIP_TRY(hipMemcpyAsync(d_buf, body + (size_t)s0 * 8, (size_t)cnt * 8,
                               hipMemcpyHostToDevice, st));
    xdr_unpack<T><<<nblocks((cnt + 1) / 2, kThreads), kThreads, 0, st>>>(
        d_buf, cnt, field + s0);
    SPS_HIP_TRY(hipStreamSynchronize(st));
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("XDR kernels failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

}  // namespace

int file_checkpoint_text(const FieldView& v, int index, const double* origin,
                         char* text, size_t capacity, size_t* length) {
  if (!length) {
    set_error("length pointer is NULL");
    return SPS_E_INVALID;
  }
  if (v.nsites > 0xffffffffull) {
    set_error("checkpoint text: fields above 2^32 sites are not supported");
    return SPS_E_UNSUPPORTED;
  }
  TextFormat fm;
  fm.header = gp_header(v.d, index, origin);
  return v.d.dtype == SPS_F32 ? write_text<float>(v, fm, text, capacity, length)
                              : write_text<double>(v, fm, text, capacity, length);
}

int file_plot_text(const FieldView& v, int format, const double* origin,
                   char* text, size_t capacity, size_t* length) {
  if (!length) {
    set_error("length pointer is NULL");
    return SPS_E_INVALID;
  }
  if (format != SPS_PLOT_GNU && format != SPS_PLOT_CSV) {
    set_error("plot text: format %d is neither SPS_PLOT_GNU nor SPS_PLOT_CSV",
              format);
    return SPS_E_INVALID;
  }
  if (v.nsites > 0xffffffffull) {
    set_error("plot text: fields above 2^32 sites are not supported");
    return SPS_E_UNSUPPORTED;
  }
  const TextFormat fm = plot_format(v.d, format, origin);
  return v.d.dtype == SPS_F32 ? write_text<float>(v, fm, text, capacity, length)
                              : write_text<double>(v, fm, text, capacity, length);
}

int file_checkpoint_read(const FieldView& v, const char* text, size_t length,
                         int* index) {
  if (!text) {
    set_error("text is NULL");
    return SPS_E_INVALID;
  }
  int idx = 0;
  const long hl = parse_header(v.d, text, length, &idx);
  if (hl < 0) return SPS_E_INVALID;
  const int rc = v.d.dtype == SPS_F32
                     ? read_text<float>(v, text + hl, (int64_t)length - hl)
                     : read_text<double>(v, text + hl, (int64_t)length - hl);
  if (rc != SPS_OK) return rc;
  if (index) *index = idx;
  return SPS_OK;
}

int file_xdr_write(const FieldView& v, int index, const double* origin,
                   void* data, size_t capacity, size_t* length) {
  if (!length) {
    set_error("length pointer is NULL");
    return SPS_E_INVALID;
  }
  const std::string header = xdr_header(v.d, index, origin);
  const size_t total = header.size() + 8 * v.nsites;
  *length = total;
  if (!data) return SPS_OK;
  if (capacity < total) {
    set_error("XDR file: capacity %zu below the %zu bytes of the file", capacity,
              total);
    return SPS_E_INVALID;
  }
  return v.d.dtype == SPS_F32 ? write_xdr<float>(v, header, (unsigned char*)data)
                              : write_xdr<double>(v, header, (unsigned char*)data);
}

int file_xdr_read(const FieldView& v, const void* data, size_t length,
                  int* index) {
  if (!data) {
    set_error("data is NULL");
    return SPS_E_INVALID;
  }
  // read_header (io/src/readxdr.cpp): counts must match the plan
  const unsigned char* b = (const unsigned char*)data;
  size_t pos = 0;
  auto get_int = [&](int* x) {
    if (pos + 4 > length) return false;
    uint32_t u = 0;
    for (int q = 0; q < 4; ++q) u = (u << 8) | b[pos++];
    *x = (int)u;
    return true;
  };
  int dim = 0, c = 0, idx = 0, dim2 = 0;
  if (!get_int(&dim) || dim != v.d.dim) {
    set_error("XDR file: dimension does not match the plan");
    return SPS_E_INVALID;
  }
  for (int a = 0; a < dim; ++a)
    if (!get_int(&c) || c != v.d.n[a]) {
      set_error("XDR file: count of axis %d does not match the plan", a);
      return SPS_E_INVALID;
    }
  pos += 16 * (size_t)dim;  // domain bounds
  if (!get_int(&idx) || !get_int(&dim2) || dim2 != dim) {
    set_error("XDR file: block header unreadable");
    return SPS_E_INVALID;
  }
  pos += 16 * (size_t)dim;  // interval bounds
  if (pos > length || length - pos < 8 * v.nsites) {
    set_error("XDR file: %zu bytes of values for a field of %zu sites",
              pos > length ? (size_t)0 : length - pos, v.nsites);
    return SPS_E_INVALID;
  }
  const int rc = v.d.dtype == SPS_F32 ? read_xdr<float>(v, b + pos)
                                      : read_xdr<double>(v, b + pos);
  if (rc != SPS_OK) return rc;
  if (index) *index = idx;
  return SPS_OK;
}

}  // namespace sps

using namespace sps;

namespace {

FieldView view_of(sps_fd_plan* p) {
  return FieldView{p->buf[p->cur], p->d, p->nsites, p->stream};
}

int plan_ready(sps_fd_plan* p) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  return SPS_OK;
}

}  // namespace

extern "C" int sps_fd_checkpoint_text(sps_fd_plan* p, int index,
                                      const double* origin, char* text,
                                      size_t capacity, size_t* length) {
  SPS_TRY(plan_ready(p));
  return file_checkpoint_text(view_of(p), index, origin, text, capacity, length);
}

extern "C" int sps_fd_plot_text(sps_fd_plan* p, int format, const double* origin,
                                char* text, size_t capacity, size_t* length) {
  SPS_TRY(plan_ready(p));
  return file_plot_text(view_of(p), format, origin, text, capacity, length);
}

extern "C" int sps_fd_checkpoint_read(sps_fd_plan* p, const char* text,
                                      size_t length, int* index) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  SPS_TRY(file_checkpoint_read(view_of(p), text, length, index));
  p->has_data = true;
  return SPS_OK;
}

extern "C" int sps_fd_checkpoint_xdr(sps_fd_plan* p, int index,
                                     const double* origin, void* data,
                                     size_t capacity, size_t* length) {
  SPS_TRY(plan_ready(p));
  return file_xdr_write(view_of(p), index, origin, data, capacity, length);
}

extern "C" int sps_fd_checkpoint_read_xdr(sps_fd_plan* p, const void* data,
                                          size_t length, int* index) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  SPS_TRY(file_xdr_read(view_of(p), data, length, index));
  p->has_data = true;
  return SPS_OK;
}
