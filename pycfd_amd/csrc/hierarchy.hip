// Aggregation-multigrid hierarchy, built on the host in native code.
//
// The algorithm is the reference's (pycfd/core/multigrid.py:324-474): per
// level the damping weight from a seeded power iteration (:129-157), the
// coarsening factors read off two diagonals of the operator (:167-204), block
// aggregates that keep every level a rectangular grid (:212-245), the 0/1
// prolongator smoothed by one damped-Jacobi pass (:248-273), the Galerkin
// coarse operator with the identity on dropped blocks (:276-287, 380-383).
//
// Every level is a grid with a compact stencil, so each sparse product is
// formed row by row with a small per-row accumulator, on a pool of host
// threads over contiguous row ranges.  Each output entry is summed by one
// thread in a fixed order (the order of the operands' rows), and the power
// iteration's norms add fixed-size chunks in index order, so the hierarchy
// does not depend on the thread count.
//
// The same pass turns every coarse level into the row dictionary the device
// V-cycle reads: stencil offsets, a per-cell class index into the level's
// distinct rows (lexicographic order), and per class 1/diag and an active flag.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <functional>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pycfd_hip.h"
#include "common.hpp"

namespace pcfd {
namespace {

constexpr int kCoarsestUnknowns = 96;     // multigrid.py:101
constexpr int kMaxLevels = 20;            // multigrid.py:104
constexpr double kMinCoarsening = 1.3;    // multigrid.py:108
constexpr double kJacobiFactor = 4.0 / 3.0;
constexpr int kPowerIterations = 12;
constexpr double kAnisotropy = 2.0;
constexpr int kNormChunk = 8192;          // fixed summation blocks of the norms
constexpr int kSerialRows = 8192;         // light loops below this run on one thread
constexpr int kSerialProducts = 64;       // product rows (long at coarse levels)

// A pool of host threads running one job over [0, n) split into contiguous
// ranges, one per thread (the caller's thread takes the first).
class Pool {
 public:
  explicit Pool(int threads) : size_(std::max(1, threads)) {
    for (int t = 1; t < size_; ++t) workers_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  int size() const { return size_; }

  // f(t, begin, end) on every non-empty range; ranges follow the thread
  // order.  Below `serial` items the caller's thread takes them all.
  void run(int64_t n, const std::function<void(int, int64_t, int64_t)>& f,
           int64_t serial = kSerialRows) {
    if (n <= 0) return;
    if (size_ == 1 || n < serial) {
      f(0, 0, n);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      job_ = &f;
      n_ = n;
      serial_ = serial;
      pending_ = size_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    range(0, f);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
    job_ = nullptr;
    if (error_) {  // the first failure of any thread, once all have finished
      std::exception_ptr e = error_;
      error_ = nullptr;
      std::rethrow_exception(e);
    }
  }
  // The range of thread t for n items.
  void bounds(int64_t n, int t, int64_t& a, int64_t& b, int64_t serial = kSerialRows) const {
    const int parts = (size_ == 1 || n < serial) ? 1 : size_;
    const int64_t chunk = (n + parts - 1) / parts;
    a = std::min<int64_t>(n, chunk * t);
    b = t < parts ? std::min<int64_t>(n, a + chunk) : a;
  }

 private:
  void range(int t, const std::function<void(int, int64_t, int64_t)>& f) {
    int64_t a, b;
    bounds(n_, t, a, b, serial_);
    if (a >= b) return;
    try {
      f(t, a, b);
    } catch (...) {
      std::lock_guard<std::mutex> g(mu_);
      if (!error_) error_ = std::current_exception();
    }
  }
  void loop(int t) {
    int seen = 0;
    for (;;) {
      const std::function<void(int, int64_t, int64_t)>* job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        job = job_;
      }
      range(t, *job);
      {
        std::lock_guard<std::mutex> g(mu_);
        --pending_;
      }
      done_.notify_one();
    }
  }

  int size_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int, int64_t, int64_t)>* job_ = nullptr;
  int64_t n_ = 0, serial_ = kSerialRows;
  int pending_ = 0, gen_ = 0;
  bool stop_ = false;
  std::exception_ptr error_;
};

// A host array left uninitialised on allocation (the build writes every
// element; zero-filling the large ones first would cost a serial pass), or a
// read-only view of the caller's memory.
template <typename T>
class Buf {
 public:
  Buf() = default;
  Buf(Buf&& o) noexcept : p_(o.p_), n_(o.n_), own_(o.own_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  Buf& operator=(Buf&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    std::swap(own_, o.own_);
    return *this;
  }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { release(); }
  void release() {
    if (own_) std::free(p_);
    p_ = nullptr;
    n_ = 0;
    own_ = true;
  }
  void alloc(int64_t n) {
    release();
    p_ = static_cast<T*>(std::malloc(sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
    if (!p_) throw std::bad_alloc();
    n_ = n;
  }
  void assign(const T* src, int64_t n) {
    alloc(n);
    if (n) std::memcpy(p_, src, sizeof(T) * (size_t)n);
  }
  void view(const T* src, int64_t n) {
    release();
    p_ = const_cast<T*>(src);
    n_ = n;
    own_ = false;
  }
  int64_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  T* data() { return p_; }
  const T* data() const { return p_; }
  T& operator[](int64_t i) { return p_[i]; }
  const T& operator[](int64_t i) const { return p_[i]; }

 private:
  T* p_ = nullptr;
  int64_t n_ = 0;
  bool own_ = true;
};

// Compressed rows, columns ascending within a row, no stored zeros.
struct Csr {
  int64_t n = 0, m = 0;
  Buf<int64_t> ptr;
  Buf<int32_t> col;
  Buf<double> val;
  int64_t nnz() const { return col.size(); }
};

// Sparse row accumulator over m columns: a touched flag and a sum per column
// (allocated lazily zeroed, so only the columns a thread touches cost
// memory) and the list of touched columns, cleared by each flush.
class RowAcc {
 public:
  explicit RowAcc(int64_t m)
      : mark_(static_cast<uint8_t*>(std::calloc(std::max<int64_t>(m, 1), 1))),
        sum_(static_cast<double*>(std::malloc(sizeof(double) * std::max<int64_t>(m, 1)))) {}
  ~RowAcc() {
    std::free(mark_);
    std::free(sum_);
  }
  RowAcc(const RowAcc&) = delete;
  RowAcc& operator=(const RowAcc&) = delete;
  bool ok() const { return mark_ && sum_; }
  void add(int32_t c, double v) {
    if (!mark_[c]) {
      mark_[c] = 1;
      sum_[c] = 0.0;
      touched_.push_back(c);
    }
    sum_[c] += v;
  }
  // Nonzero sums in ascending column order appended to (col, val); cleared.
  void flush(std::vector<int32_t>& col, std::vector<double>& val) {
    std::sort(touched_.begin(), touched_.end());
    for (int32_t c : touched_) {
      if (sum_[c] != 0.0) {
        col.push_back(c);
        val.push_back(sum_[c]);
      }
      mark_[c] = 0;
    }
    touched_.clear();
  }

 private:
  uint8_t* mark_;
  double* sum_;
  std::vector<int32_t> touched_;
};

// Rows [0, n) of a product formed by row(i, acc) on the pool: each thread
// appends its rows to its own buffers, which are then placed by row offset.
template <class RowFn>
Csr build_rows(Pool& pool, int64_t n, int64_t m, RowFn row) {
  const int T = pool.size();
  std::vector<std::vector<int32_t>> cols(T);
  std::vector<std::vector<double>> vals(T);
  std::vector<int64_t> len(n + 1, 0);
  std::atomic<bool> oom{false};
  pool.run(
      n,
      [&](int t, int64_t a, int64_t b) {
    RowAcc acc(m);
    if (!acc.ok()) {
      oom = true;
      return;
    }
    // local buffers: the shared vectors' headers would be written by every
    // thread on each append (one cache line for all of them)
    std::vector<int32_t> c;
    std::vector<double> v;
    for (int64_t i = a; i < b; ++i) {
      const size_t before = c.size();
      row(i, acc);
      acc.flush(c, v);
      len[i + 1] = (int64_t)(c.size() - before);
    }
    cols[t] = std::move(c);
    vals[t] = std::move(v);
      },
      kSerialProducts);
  if (oom) throw std::bad_alloc();
  Csr out;
  out.n = n;
  out.m = m;
  out.ptr.alloc(n + 1);
  out.ptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) out.ptr[i + 1] = out.ptr[i] + len[i + 1];
  out.col.alloc(out.ptr[n]);
  out.val.alloc(out.ptr[n]);
  pool.run(
      n,
      [&](int t, int64_t a, int64_t) {
        std::copy(cols[t].begin(), cols[t].end(), out.col.data() + out.ptr[a]);
        std::copy(vals[t].begin(), vals[t].end(), out.val.data() + out.ptr[a]);
      },
      kSerialProducts);
  return out;
}

// sqrt(sum x^2): fixed chunks summed in index order, then the chunk sums.
double norm2(Pool& pool, const std::vector<double>& x) {
  const int64_t n = (int64_t)x.size();
  const int64_t nch = (n + kNormChunk - 1) / kNormChunk;
  std::vector<double> part(nch, 0.0);
  pool.run(
      nch,
      [&](int, int64_t a, int64_t b) {
        for (int64_t c = a; c < b; ++c) {
          double s = 0.0;
          const int64_t e = std::min<int64_t>(n, (c + 1) * kNormChunk);
          for (int64_t i = c * kNormChunk; i < e; ++i) s += x[i] * x[i];
          part[c] = s;
        }
      },
      8);
  double s = 0.0;
  for (double v : part) s += v;
  return std::sqrt(s);
}

// rho(D^-1 A) by power iteration from the seeded start vector x0, nudged up
// 5% (multigrid.py:129-157).
template <class Apply>
double spectral_radius(Pool& pool, int64_t n, const std::vector<double>& dinv, const double* x0,
                       Apply&& apply_row) {
  if (n == 0) return 1.0;
  std::vector<double> x(x0, x0 + n), y(n);
  double nrm = norm2(pool, x);
  if (nrm == 0.0) return 1.0;
  for (double& v : x) v /= nrm;
  double rho = 1.0;
  for (int it = 0; it < kPowerIterations; ++it) {
    pool.run(n, [&](int, int64_t a, int64_t b) {
      for (int64_t i = a; i < b; ++i) y[i] = dinv[i] * apply_row(i, x.data());
    });
    nrm = norm2(pool, y);
    if (nrm == 0.0) return rho;
    rho = nrm;
    pool.run(n, [&](int, int64_t a, int64_t b) {
      for (int64_t i = a; i < b; ++i) x[i] = y[i] / nrm;
    });
  }
  return rho * 1.05;
}

double entry(const Csr& A, int64_t i, int64_t j) {
  double s = 0.0;
  for (int64_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e)
    if (A.col[e] == j) s += A.val[e];
  return s;
}

// Mean |A[i, i + off]| over i < n - off (multigrid.py:193-197), summed in
// fixed chunks in index order.
double strength(Pool& pool, const Csr& A, int64_t off) {
  if (off >= A.n) return 0.0;
  const int64_t cnt = A.n - off;
  const int64_t nch = (cnt + kNormChunk - 1) / kNormChunk;
  std::vector<double> part(nch, 0.0);
  pool.run(
      nch,
      [&](int, int64_t a, int64_t b) {
        for (int64_t c = a; c < b; ++c) {
          double s = 0.0;
          const int64_t e = std::min<int64_t>(cnt, (c + 1) * kNormChunk);
          for (int64_t i = c * kNormChunk; i < e; ++i) s += std::fabs(entry(A, i, i + off));
          part[c] = s;
        }
      },
      8);
  double s = 0.0;
  for (double v : part) s += v;
  return s / (double)cnt;
}

struct Level {
  int nx = 0, ny = 0;
  Csr A;
  std::vector<double> dinv;
  double weight = 1.0;
  std::vector<uint8_t> active;
  int fx = 1, fy = 1;  // towards the next level
  int pin = 0;
  int rx = 0, ry = 0;  // stencil radii of the operator
  int64_t nnz0 = -1;   // level 0: entries of the caller's operator (not kept)
  // row dictionary; cls is empty when a level has too many distinct rows
  // for its classes to pay (the fine level of a stretched mesh)
  std::vector<int8_t> di, dj;
  std::vector<uint32_t> cls;
  std::vector<double> table, cdinv;
  std::vector<uint8_t> cact;
};


// Division of non-negative 32-bit values by a fixed divisor as a multiply
// and shift (exact for every numerator below 2^32).
struct FastDiv {
  uint64_t d, m;
  explicit FastDiv(uint64_t div) : d(div), m(~0ull / div + 1) {}
  uint64_t div(uint64_t a) const { return (uint64_t)(((unsigned __int128)a * m) >> 64); }
};

inline int64_t floor_mod(int64_t a, int64_t m) {
  const int64_t r = a % m;
  return r < 0 ? r + m : r;
}

// One level's operator as a row dictionary: the stencil offsets (and radii)
// always, the classes unless more than max_classes distinct rows turn up
// (max_classes < 0: no limit).
int build_dictionary(Pool& pool, Level& L, bool px, bool py, int64_t max_classes) {
  const int64_t n = L.A.n;
  const int nx = L.nx, ny = L.ny;
  const FastDiv by_ny((uint64_t)ny);
  auto offset = [&](int64_t r, int32_t c, int& di, int& dj) {
    const int64_t ri = (int64_t)by_ny.div((uint64_t)r), rj = r - ri * ny;
    const int64_t ci = (int64_t)by_ny.div((uint64_t)c), cj = c - ci * ny;
    int64_t a = ci - ri, b = cj - rj;
    if (px && nx > 1) a = floor_mod(a + nx / 2, nx) - nx / 2;
    if (py && ny > 1) b = floor_mod(b + ny / 2, ny) - ny / 2;
    di = (int)a;
    dj = (int)b;
  };
  // distinct offsets in (di, dj) order, from per-thread presence tables
  const int T = pool.size();
  std::vector<std::vector<uint8_t>> seen(T);
  std::vector<uint8_t> bad(T, 0);
  pool.run(n, [&](int t, int64_t a, int64_t b) {
    seen[t].assign(256 * 256, 0);
    for (int64_t r = a; r < b; ++r)
      for (int64_t e = L.A.ptr[r]; e < L.A.ptr[r + 1]; ++e) {
        int di, dj;
        offset(r, L.A.col[e], di, dj);
        if (di < -127 || di > 127 || dj < -127 || dj > 127) {
          bad[t] = 1;
          continue;
        }
        seen[t][(di + 128) * 256 + (dj + 128)] = 1;
      }
  });
  std::vector<uint8_t> present(256 * 256, 0);
  bool range_ok = true;
  for (int t = 0; t < T; ++t) {
    range_ok = range_ok && !bad[t];
    if (seen[t].empty()) continue;
    for (int k = 0; k < 256 * 256; ++k) present[k] |= seen[t][k];
  }
  if (!range_ok) {
    set_last_error("coarse stencil offset exceeds the int8 range");
    return kErrUnsupported;
  }
  std::vector<int> slot(256 * 256, -1);
  int S = 0, centre = -1;
  L.di.clear();
  L.dj.clear();
  L.rx = L.ry = 0;
  for (int k = 0; k < 256 * 256; ++k)
    if (present[k]) {
      slot[k] = S++;
      L.di.push_back((int8_t)(k / 256 - 128));
      L.dj.push_back((int8_t)(k % 256 - 128));
      L.rx = std::max(L.rx, std::abs(k / 256 - 128));
      L.ry = std::max(L.ry, std::abs(k % 256 - 128));
      if (k / 256 == 128 && k % 256 == 128) centre = S - 1;
    }
  if (centre < 0) {
    set_last_error("a coarse operator without diagonal entries");
    return kErrBadArg;
  }
  const int W = S + 1;  // coefficients, then the active flag
  auto dense = [&](int64_t r, double* row) {
    std::fill(row, row + W, 0.0);
    for (int64_t e = L.A.ptr[r]; e < L.A.ptr[r + 1]; ++e) {
      int di, dj;
      offset(r, L.A.col[e], di, dj);
      const double v = L.A.val[e];
      row[slot[(di + 128) * 256 + (dj + 128)]] = v == 0.0 ? 0.0 : v;
    }
    row[S] = L.active[r] ? 1.0 : 0.0;
  };
  auto hash = [&](const double* row) {
    uint64_t h = 0x9e3779b97f4a7c15ull;
    for (int k = 0; k < W; ++k) {
      uint64_t b;
      std::memcpy(&b, row + k, 8);
      h ^= b + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
      h *= 0xff51afd7ed558ccdull;
    }
    return h ^ (h >> 33);
  };
  // Distinct rows: per thread, then merged in thread order.  A map entry
  // lists the rows sharing one hash, compared bit for bit.
  struct Distinct {
    std::unordered_map<uint64_t, std::vector<int64_t>> by_hash;  // hash -> rows
    std::vector<double> rows;                                    // W per distinct row
  };
  auto find = [&](const Distinct& D, uint64_t h, const double* row) -> int64_t {
    auto it = D.by_hash.find(h);
    if (it == D.by_hash.end()) return -1;
    for (int64_t k : it->second)
      if (std::memcmp(D.rows.data() + k * W, row, sizeof(double) * W) == 0) return k;
    return -1;
  };
  auto insert = [&](Distinct& D, uint64_t h, const double* row) {
    const int64_t k = (int64_t)(D.rows.size() / W);
    D.rows.insert(D.rows.end(), row, row + W);
    D.by_hash[h].push_back(k);
  };
  std::vector<Distinct> local(T);
  std::vector<int32_t> slot_of(n);  // each row's distinct row in its thread's list
  std::atomic<bool> over{false};
  pool.run(n, [&](int t, int64_t a, int64_t b) {
    std::vector<double> row(W);
    for (int64_t r = a; r < b && !over; ++r) {
      dense(r, row.data());
      const uint64_t h = hash(row.data());
      int64_t k = find(local[t], h, row.data());
      if (k < 0) {
        k = (int64_t)(local[t].rows.size() / W);
        insert(local[t], h, row.data());
        if (max_classes >= 0 && k + 1 > max_classes) over = true;
      }
      slot_of[r] = (int32_t)k;
    }
  });
  L.cls.clear();
  L.table.clear();
  L.cdinv.clear();
  L.cact.clear();
  if (over) return kOk;
  Distinct all;
  std::vector<std::vector<int64_t>> to_all(T);  // thread's distinct row -> merged
  for (int t = 0; t < T; ++t) {
    const int64_t m = (int64_t)(local[t].rows.size() / W);
    to_all[t].resize(m);
    for (int64_t k = 0; k < m; ++k) {
      const double* row = local[t].rows.data() + k * W;
      const uint64_t h = hash(row);
      int64_t g = find(all, h, row);
      if (g < 0) {
        g = (int64_t)(all.rows.size() / W);
        insert(all, h, row);
      }
      to_all[t][k] = g;
    }
    local[t] = Distinct{};
  }
  const int64_t nc = (int64_t)(all.rows.size() / W);
  if (max_classes >= 0 && nc > max_classes) return kOk;
  if (nc >= (int64_t)1 << 32) {
    set_last_error("%lld distinct operator rows exceed the class index range", (long long)nc);
    return kErrUnsupported;
  }
  // lexicographic order of the distinct rows (coefficients, then the flag)
  std::vector<int64_t> order(nc);
  for (int64_t k = 0; k < nc; ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    const double* ra = all.rows.data() + a * W;
    const double* rb = all.rows.data() + b * W;
    for (int k = 0; k < W; ++k) {
      if (ra[k] < rb[k]) return true;
      if (ra[k] > rb[k]) return false;
    }
    return false;
  });
  std::vector<uint32_t> rank(nc);
  for (int64_t k = 0; k < nc; ++k) rank[order[k]] = (uint32_t)k;
  L.table.assign((size_t)nc * S, 0.0);
  L.cdinv.assign(nc, 0.0);
  L.cact.assign(nc, 0);
  for (int64_t k = 0; k < nc; ++k) {
    const double* row = all.rows.data() + order[k] * W;
    std::copy(row, row + S, L.table.begin() + k * S);
    const double d = row[centre];
    L.cdinv[k] = 1.0 / (d == 0.0 ? 1.0 : d);
    L.cact[k] = row[S] != 0.0;
  }
  L.cls.resize(n);
  pool.run(n, [&](int t, int64_t a, int64_t b) {  // the ranges of the first pass
    for (int64_t r = a; r < b; ++r) L.cls[r] = rank[to_all[t][slot_of[r]]];
  });
  return kOk;
}

// 1/diag and the damping weight of a level (multigrid.py:354-362).  The
// power iteration applies the operator from the row dictionary when the
// level has one (a class index per cell instead of a compressed row), in the
// dictionary's offset order.
void make_level(Pool& pool, Level& L, bool px, bool py, const double* x0) {
  const int64_t n = L.A.n;
  L.dinv.resize(n);
  pool.run(n, [&](int, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      const double d = entry(L.A, i, i);
      L.dinv[i] = 1.0 / (d == 0.0 ? 1.0 : d);
    }
  });
  double rho;
  if (L.cls.empty()) {
    rho = spectral_radius(pool, n, L.dinv, x0, [&](int64_t i, const double* x) {
      double s = 0.0;
      for (int64_t e = L.A.ptr[i]; e < L.A.ptr[i + 1]; ++e) s += L.A.val[e] * x[L.A.col[e]];
      return s;
    });
  } else {
    const int S = (int)L.di.size(), nx = L.nx, ny = L.ny;
    std::vector<int64_t> lin(S);
    for (int s = 0; s < S; ++s) lin[s] = (int64_t)L.di[s] * ny + L.dj[s];
    const FastDiv by_ny((uint64_t)ny);
    rho = spectral_radius(pool, n, L.dinv, x0, [&](int64_t i, const double* x) {
      const double* row = L.table.data() + (size_t)L.cls[i] * S;
      const int64_t xi = (int64_t)by_ny.div((uint64_t)i), yi = i - xi * ny;
      double acc = 0.0;
      if (xi >= L.rx && xi < nx - L.rx && yi >= L.ry && yi < ny - L.ry) {
        for (int s = 0; s < S; ++s)
          if (row[s] != 0.0) acc += row[s] * x[i + lin[s]];
        return acc;
      }
      for (int s = 0; s < S; ++s) {
        if (row[s] == 0.0) continue;
        int64_t a = xi + L.di[s], b = yi + L.dj[s];
        if (a < 0 || a >= nx) {
          if (!px) continue;
          a = floor_mod(a, nx);
        }
        if (b < 0 || b >= ny) {
          if (!py) continue;
          b = floor_mod(b, ny);
        }
        acc += row[s] * x[a * ny + b];
      }
      return acc;
    });
  }
  L.weight = rho > 0.0 ? kJacobiFactor / rho : 1.0;
}

// Coarse row I depends on the fine level only through the block of I grown
// by three stencil radii: the operator rows within two radii (the
// prolongator rows P(j, :) that A P combines, each from its own operator
// row) and the active cells within three (the aggregates those prolongator
// rows reach).  Where that window lies inside the grid and its cells carry
// the same classes in the same places, two coarse rows are formed by the
// same operations on the same values, so they are equal entry for entry up
// to the translation.  The Galerkin product is therefore formed once per
// distinct window (the representative, its first coarse row) and translated
// to the others.  A window that leaves the grid across a periodic axis is
// its own class.
struct Blocks {
  int nx, ny, fx, fy, ncx, ncy;
  bool px, py;
  // fine cells of the block of coarse cell I grown by (gx, gy)
  void window(int64_t I, int gx, int gy, int64_t& x0, int64_t& x1, int64_t& y0,
              int64_t& y1) const {
    const int64_t ix = I / ncy, iy = I - ix * ncy;
    x0 = ix * fx - gx;
    x1 = std::min<int64_t>((ix + 1) * fx, nx) + gx;
    y0 = iy * fy - gy;
    y1 = std::min<int64_t>((iy + 1) * fy, ny) + gy;
  }
  bool wraps(int64_t x0, int64_t x1, int64_t y0, int64_t y1) const {
    return (px && (x0 < 0 || x1 > nx)) || (py && (y0 < 0 || y1 > ny));
  }
  // the grid cell of window position (x, y), or -1 outside a bounded axis
  int64_t cell(int64_t x, int64_t y) const {
    if (x < 0 || x >= nx) {
      if (!px) return -1;
      x = floor_mod(x, nx);
    }
    if (y < 0 || y >= ny) {
      if (!py) return -1;
      y = floor_mod(y, ny);
    }
    return x * ny + y;
  }
};

// The class window of coarse row I into w (row-major; ~0u outside the grid).
// False when the window wraps a periodic axis.
bool class_window(const Blocks& B, const Level& F, int64_t I, std::vector<uint32_t>& w) {
  int64_t x0, x1, y0, y1;
  B.window(I, 3 * F.rx, 3 * F.ry, x0, x1, y0, y1);
  if (B.wraps(x0, x1, y0, y1)) return false;
  w.clear();
  if (x0 >= 0 && x1 <= B.nx && y0 >= 0 && y1 <= B.ny) {
    for (int64_t x = x0; x < x1; ++x) {
      const uint32_t* src = F.cls.data() + x * B.ny;
      w.insert(w.end(), src + y0, src + y1);
    }
    return true;
  }
  for (int64_t x = x0; x < x1; ++x)
    for (int64_t y = y0; y < y1; ++y) {
      const int64_t k = B.cell(x, y);
      w.push_back(k < 0 ? ~0u : F.cls[k]);
    }
  return true;
}

uint64_t hash_words(const std::vector<uint32_t>& w) {
  uint64_t h = 0xcbf29ce484222325ull ^ (uint64_t)w.size();
  for (uint32_t v : w) {
    h ^= v;
    h *= 0x100000001b3ull;
    h ^= h >> 29;
  }
  return h;
}

// The representative of every coarse row (itself for the first row of each
// distinct window and for windows that wrap).  Windows are compared in full
// (the hash only selects candidates).
std::vector<int64_t> representatives(Pool& pool, const Blocks& B, const Level& F, int64_t nc) {
  std::vector<int64_t> rep(nc);
  for (int64_t I = 0; I < nc; ++I) rep[I] = I;
  if (F.cls.empty()) return rep;  // no classes: every row on its own
  struct Groups {
    std::unordered_map<uint64_t, std::vector<int64_t>> by_hash;  // hash -> slots
    std::vector<int64_t> first;                                  // first row per slot
    std::vector<std::vector<uint32_t>> window;                   // its window
    int64_t find(uint64_t h, const std::vector<uint32_t>& w) const {
      auto it = by_hash.find(h);
      if (it == by_hash.end()) return -1;
      for (int64_t k : it->second)
        if (window[k] == w) return k;
      return -1;
    }
    void add(uint64_t h, int64_t I, const std::vector<uint32_t>& w) {
      by_hash[h].push_back((int64_t)first.size());
      first.push_back(I);
      window.push_back(w);
    }
  };
  const int T = pool.size();
  std::vector<Groups> local(T);
  std::vector<uint64_t> sig(nc, 0);
  std::vector<int32_t> slot_of(nc, -1);  // -1: wraps, its own class
  pool.run(nc, [&](int t, int64_t a, int64_t b) {
    std::vector<uint32_t> w;
    for (int64_t I = a; I < b; ++I) {
      if (!class_window(B, F, I, w)) continue;
      sig[I] = hash_words(w);
      int64_t k = local[t].find(sig[I], w);
      if (k < 0) {
        k = (int64_t)local[t].first.size();
        local[t].add(sig[I], I, w);
      }
      slot_of[I] = (int32_t)k;
    }
  });
  Groups all;  // merged in thread order: each class's first row overall
  std::vector<std::vector<int64_t>> to_all(T);
  for (int t = 0; t < T; ++t) {
    to_all[t].resize(local[t].first.size());
    for (size_t k = 0; k < local[t].first.size(); ++k) {
      const int64_t I = local[t].first[k];
      int64_t g = all.find(sig[I], local[t].window[k]);
      if (g < 0) {
        g = (int64_t)all.first.size();
        all.add(sig[I], I, local[t].window[k]);
      }
      to_all[t][k] = g;
    }
    local[t] = Groups{};
  }
  pool.run(nc, [&](int t, int64_t a, int64_t b) {  // the ranges of the first pass
    for (int64_t I = a; I < b; ++I)
      if (slot_of[I] >= 0) rep[I] = all.first[to_all[t][slot_of[I]]];
  });
  return rep;
}

// Rows of a product formed only for the listed rows: row k of the result is
// row list[k]; pos[i] is i's place in the list or -1.
struct Listed {
  std::vector<int64_t> list;
  std::vector<int64_t> pos;
  void mark_done(int64_t n) {
    pos.assign(n, -1);
    for (size_t k = 0; k < list.size(); ++k) pos[list[k]] = (int64_t)k;
  }
};

// Coarsens F into C (multigrid.py:370-383): factors, aggregates, the smoothed
// prolongator and the Galerkin operator with the identity on dropped blocks.
// False when the level would not shrink enough; C is then untouched.
bool coarsen(Pool& pool, Level& F, Level& C, bool px, bool py, bool every_row) {
  const int nx = F.nx, ny = F.ny;
  const int64_t nf = F.A.n;
  int fx, fy;  // multigrid.py:186-204
  if (nx < 2 && ny < 2) {
    fx = fy = 1;
  } else {
    const double sx = strength(pool, F.A, ny), sy = strength(pool, F.A, 1);
    if (nx >= 2 && sx > kAnisotropy * sy) {
      fx = 2;
      fy = 1;
    } else if (ny >= 2 && sy > kAnisotropy * sx) {
      fx = 1;
      fy = 2;
    } else {
      fx = nx >= 2 ? 2 : 1;
      fy = ny >= 2 ? 2 : 1;
    }
  }
  const int ncx = std::max((nx + fx - 1) / fx, 1), ncy = std::max((ny + fy - 1) / fy, 1);
  const int64_t nc = (int64_t)ncx * ncy;
  if (nc == 0 || (double)nf < kMinCoarsening * (double)nc) return false;
  // aggregates (multigrid.py:212-245)
  std::vector<int32_t> label(nf);
  std::vector<uint8_t> cact(nc, 0);
  const FastDiv by_ny((uint64_t)ny);
  pool.run(nf, [&](int, int64_t a, int64_t b) {
    for (int64_t k = a; k < b; ++k) {
      const int64_t i = (int64_t)by_ny.div((uint64_t)k), j = k - i * ny;
      label[k] = F.active[k] ? (int32_t)((i / fx) * ncy + j / fy) : -1;
    }
  });
  for (int64_t k = 0; k < nf; ++k)  // one writer per block
    if (label[k] >= 0) cact[label[k]] = 1;
  const Blocks B{nx, ny, fx, fy, ncx, ncy, px, py};
  std::vector<int64_t> rep;
  if (every_row) {
    rep.resize(nc);
    for (int64_t I = 0; I < nc; ++I) rep[I] = I;
  } else {
    rep = representatives(pool, B, F, nc);
  }
  // fine rows the representatives need: P within two radii of their
  // blocks, A P within one
  std::vector<uint8_t> need(nf, 0);  // 1: P, 2: P and A P
  Listed reps;
  for (int64_t I = 0; I < nc; ++I) {
    if (rep[I] != I) continue;
    reps.list.push_back(I);
    for (int g = 2; g >= 1; --g) {
      int64_t x0, x1, y0, y1;
      B.window(I, g * F.rx, g * F.ry, x0, x1, y0, y1);
      for (int64_t x = x0; x < x1; ++x)
        for (int64_t y = y0; y < y1; ++y) {
          const int64_t k = B.cell(x, y);
          if (k >= 0) need[k] = (uint8_t)std::max<int>(need[k], 3 - g);
        }
    }
  }
  reps.mark_done(nc);
  Listed pr, apr;  // rows of P, rows of A P
  for (int64_t k = 0; k < nf; ++k) {
    if (need[k]) pr.list.push_back(k);
    if (need[k] == 2) apr.list.push_back(k);
  }
  pr.mark_done(nf);
  apr.mark_done(nf);
  // P = T - diag(w D^-1) (A T), T the 0/1 aggregation (multigrid.py:248-273)
  const double w = F.weight;
  Csr P = build_rows(pool, (int64_t)pr.list.size(), nc, [&](int64_t q, RowAcc& acc) {
    const int64_t i = pr.list[q];
    thread_local std::vector<int32_t> tc;
    thread_local std::vector<double> tv;
    tc.clear();
    tv.clear();
    for (int64_t e = F.A.ptr[i]; e < F.A.ptr[i + 1]; ++e) {
      const int32_t c = label[F.A.col[e]];
      if (c >= 0) acc.add(c, F.A.val[e]);
    }
    acc.flush(tc, tv);  // (A T)(i, :), zero sums dropped
    const double wd = w * F.dinv[i];
    bool own = false;
    for (size_t k = 0; k < tc.size(); ++k) {
      const double x = wd * tv[k];
      if (tc[k] == label[i]) {
        acc.add(tc[k], 1.0 - x);
        own = true;
      } else {
        acc.add(tc[k], 0.0 - x);
      }
    }
    if (label[i] >= 0 && !own) acc.add(label[i], 1.0);
  });
  // A P on the rows next to the representatives' blocks
  Csr AP = build_rows(pool, (int64_t)apr.list.size(), nc, [&](int64_t q, RowAcc& acc) {
    const int64_t i = apr.list[q];
    for (int64_t e = F.A.ptr[i]; e < F.A.ptr[i + 1]; ++e) {
      const int64_t j = pr.pos[F.A.col[e]];
      const double a = F.A.val[e];
      for (int64_t k = P.ptr[j]; k < P.ptr[j + 1]; ++k) acc.add(P.col[k], a * P.val[k]);
    }
  });
  // P^T (A P) on the representatives: rows i with P(i, I) != 0 in ascending
  // order, then the identity on dropped blocks (multigrid.py:276-287)
  Csr R = build_rows(pool, (int64_t)reps.list.size(), nc, [&](int64_t q, RowAcc& acc) {
    const int64_t I = reps.list[q];
    thread_local std::vector<int64_t> rows;
    rows.clear();
    int64_t x0, x1, y0, y1;
    B.window(I, F.rx, F.ry, x0, x1, y0, y1);
    for (int64_t x = x0; x < x1; ++x)
      for (int64_t y = y0; y < y1; ++y) {
        const int64_t k = B.cell(x, y);
        if (k >= 0) rows.push_back(k);
      }
    if (B.wraps(x0, x1, y0, y1)) {
      std::sort(rows.begin(), rows.end());
      rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    }
    for (int64_t i : rows) {
      const int64_t pi = pr.pos[i], ai = apr.pos[i];
      const int32_t* c0 = P.col.data() + P.ptr[pi];
      const int32_t* c1 = P.col.data() + P.ptr[pi + 1];
      const int32_t* hit = std::lower_bound(c0, c1, (int32_t)I);
      if (hit == c1 || *hit != (int32_t)I) continue;
      const double p = P.val[P.ptr[pi] + (hit - c0)];
      for (int64_t e = AP.ptr[ai]; e < AP.ptr[ai + 1]; ++e) acc.add(AP.col[e], p * AP.val[e]);
    }
    if (!cact[I]) acc.add((int32_t)I, 1.0);
  });
  // every row from its representative, translated
  C.nx = ncx;
  C.ny = ncy;
  C.active = cact;
  C.A.n = C.A.m = nc;
  C.A.ptr.alloc(nc + 1);
  C.A.ptr[0] = 0;
  for (int64_t I = 0; I < nc; ++I) {
    const int64_t r = reps.pos[rep[I]];
    C.A.ptr[I + 1] = C.A.ptr[I] + (R.ptr[r + 1] - R.ptr[r]);
  }
  C.A.col.alloc(C.A.ptr[nc]);
  C.A.val.alloc(C.A.ptr[nc]);
  pool.run(nc, [&](int, int64_t a, int64_t b) {
    for (int64_t I = a; I < b; ++I) {
      const int64_t r = reps.pos[rep[I]];
      const int64_t shift = I - rep[I];  // translations keep the window inside the grid
      int64_t o = C.A.ptr[I];
      for (int64_t e = R.ptr[r]; e < R.ptr[r + 1]; ++e, ++o) {
        C.A.col[o] = (int32_t)(R.col[e] + shift);
        C.A.val[o] = R.val[e];
      }
    }
  });
  F.fx = fx;
  F.fy = fy;
  return true;
}

}  // namespace

struct MgSetup {
  std::vector<Level> levels;
  bool px = false, py = false, singular = false;
};

}  // namespace pcfd

extern "C" {

int pcfd_mg_build(const pcfd_mg_operator* op, pcfd_mg_setup** out) {
  using namespace pcfd;
  if (!op || !out || !op->indptr || !op->indices || !op->data || !op->active || !op->x0 ||
      op->nx < 1 || op->ny < 1) {
    set_last_error("pcfd_mg_build: bad argument");
    return kErrBadArg;
  }
  *out = nullptr;
  const int64_t n = (int64_t)op->nx * op->ny;
  if (n >= ((int64_t)1 << 31)) {
    set_last_error("pcfd_mg_build: %lld cells exceed 32-bit columns", (long long)n);
    return kErrUnsupported;
  }
  const int64_t nnz = op->indptr[n];
  if (op->indptr[0] != 0 || nnz < 0) {
    set_last_error("pcfd_mg_build: indptr must start at 0");
    return kErrBadArg;
  }
  for (int64_t i = 0; i < n; ++i)
    if (op->indptr[i + 1] < op->indptr[i]) {
      set_last_error("pcfd_mg_build: indptr decreases at row %lld", (long long)i);
      return kErrBadArg;
    }
  for (int64_t e = 0; e < nnz; ++e)
    if (op->indices[e] < 0 || op->indices[e] >= n) {
      set_last_error("pcfd_mg_build: column %d out of range", op->indices[e]);
      return kErrBadArg;
    }
  auto* H = new MgSetup;
  H->px = op->periodic_x != 0;
  H->py = op->periodic_y != 0;
  H->singular = op->singular != 0;
  try {
    Pool pool(std::max(1, std::min(op->threads, 64)));
    Level L;
    L.nx = op->nx;
    L.ny = op->ny;
    L.A.n = L.A.m = n;
    // the caller's arrays serve as level 0's operator during the build,
    // unless a row's columns are out of order (each coarse row's sums run in
    // the order of its window): then a sorted copy
    std::atomic<bool> unsorted{false};
    pool.run(n, [&](int, int64_t a, int64_t b) {
      for (int64_t i = a; i < b && !unsorted; ++i)
        if (!std::is_sorted(op->indices + op->indptr[i], op->indices + op->indptr[i + 1]))
          unsorted = true;
    });
    if (!unsorted) {
      L.A.ptr.view(op->indptr, n + 1);
      L.A.col.view(op->indices, nnz);
      L.A.val.view(op->data, nnz);
    } else {
      L.A.ptr.assign(op->indptr, n + 1);
      L.A.col.assign(op->indices, nnz);
      L.A.val.assign(op->data, nnz);
      pool.run(n, [&](int, int64_t a, int64_t b) {
        std::vector<std::pair<int32_t, double>> row;
        for (int64_t i = a; i < b; ++i) {
          const int64_t e0 = L.A.ptr[i], e1 = L.A.ptr[i + 1];
          row.clear();
          for (int64_t e = e0; e < e1; ++e) row.push_back(std::make_pair(L.A.col[e], L.A.val[e]));
          std::stable_sort(row.begin(), row.end(),
                           [](const std::pair<int32_t, double>& x,
                              const std::pair<int32_t, double>& y) { return x.first < y.first; });
          for (int64_t e = e0; e < e1; ++e) {
            L.A.col[e] = row[e - e0].first;
            L.A.val[e] = row[e - e0].second;
          }
        }
      });
    }
    L.active.assign(op->active, op->active + n);
    for (auto& a : L.active) a = a ? 1 : 0;
    // the fine level's classes pay only when its rows repeat (a stretched
    // mesh has as many as cells); coarse levels need theirs for the device
    if (int rc = build_dictionary(pool, L, H->px, H->py, n / 8)) {
      delete H;
      return rc;
    }
    make_level(pool, L, H->px, H->py, op->x0);
    H->levels.push_back(std::move(L));
    const int64_t stop = std::max<int64_t>(kCoarsestUnknowns, op->coarse_max);
    while ((int)H->levels.size() < kMaxLevels && H->levels.back().A.n > stop) {
      Level C;
      if (!coarsen(pool, H->levels.back(), C, H->px, H->py, op->every_row != 0)) break;
      if (int rc = build_dictionary(pool, C, H->px, H->py, -1)) {
        delete H;
        return rc;
      }
      make_level(pool, C, H->px, H->py, op->x0);
      H->levels.push_back(std::move(C));
    }
  } catch (const std::bad_alloc&) {
    delete H;
    set_last_error("pcfd_mg_build: out of host memory");
    return kErrBadArg;
  } catch (const std::exception& e) {
    delete H;
    set_last_error("pcfd_mg_build: %s", e.what());
    return kErrBadArg;
  }
  // level 0's operator is the caller's: not kept past the call
  H->levels[0].nnz0 = H->levels[0].A.nnz();
  H->levels[0].A = Csr{};
  Level& last = H->levels.back();
  last.pin = 0;
  for (size_t k = 0; k < last.active.size(); ++k)
    if (last.active[k]) {
      last.pin = (int)k;
      break;
    }
  *out = reinterpret_cast<pcfd_mg_setup*>(H);
  return kOk;
}

int pcfd_mg_free(pcfd_mg_setup* h) {
  delete reinterpret_cast<pcfd::MgSetup*>(h);
  return pcfd::kOk;
}

int pcfd_mg_level_info(const pcfd_mg_setup* h, int32_t level, pcfd_mg_built_level* info) {
  using namespace pcfd;
  const auto* H = reinterpret_cast<const MgSetup*>(h);
  if (!H || !info) {
    set_last_error("pcfd_mg_level_info: null argument");
    return kErrBadArg;
  }
  if (level < 0 || level >= (int)H->levels.size()) {
    set_last_error("pcfd_mg_level_info: level %d of %d", level, (int)H->levels.size());
    return kErrBadArg;
  }
  const Level& L = H->levels[level];
  info->n_levels = (int32_t)H->levels.size();
  info->nx = L.nx;
  info->ny = L.ny;
  info->fx = L.fx;
  info->fy = L.fy;
  info->nnz = L.nnz0 >= 0 ? L.nnz0 : L.A.nnz();
  info->weight = L.weight;
  info->pin = L.pin;
  info->n_offsets = (int32_t)L.di.size();
  info->n_classes = (int64_t)L.cdinv.size();
  return kOk;
}

int pcfd_mg_level_copy(const pcfd_mg_setup* h, int32_t level, int64_t* indptr, int32_t* indices,
                       double* data, double* dinv, uint8_t* active, int8_t* di, int8_t* dj,
                       uint32_t* cls, double* table, double* class_dinv, uint8_t* class_active) {
  using namespace pcfd;
  const auto* H = reinterpret_cast<const MgSetup*>(h);
  if (!H || level < 0 || level >= (int)H->levels.size()) {
    set_last_error("pcfd_mg_level_copy: bad level");
    return kErrBadArg;
  }
  const Level& L = H->levels[level];
  auto put = [](auto* dst, const auto& src) {
    if (dst && !src.empty()) std::memcpy(dst, src.data(), (size_t)src.size() * sizeof(src[0]));
  };
  put(indptr, L.A.ptr);
  put(indices, L.A.col);
  put(data, L.A.val);
  put(dinv, L.dinv);
  put(active, L.active);
  put(di, L.di);
  put(dj, L.dj);
  put(cls, L.cls);
  put(table, L.table);
  put(class_dinv, L.cdinv);
  put(class_active, L.cact);
  return kOk;
}

}  // extern "C"
