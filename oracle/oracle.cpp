// CPU oracle for the pattern-evaluator fit -- TEST INFRASTRUCTURE ONLY.
//
// A plain, deliberately simple CPU restatement of the reference algorithms of
// dominichofer/Cassandra.  Used by the tests to check the HIP library and by
// bench.py as the timed CPU baseline.  The product library never uses it.
// Functions cite the reference file:line they restate (paths relative to the
// reference repository root).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- squares --
// Bit b of a bitboard is the square (x, y) with b = x + 8y (Board/Field.h:7-18,
// Pattern/Names.h:6-19).  Symmetries are written as square maps.
inline int sq(int x, int y) { return x + 8 * y; }

template <class F>
uint64_t map_squares(uint64_t b, F f)
{
    uint64_t out = 0;
    for (int s = 0; s < 64; s++)
        if ((b >> s) & 1) {
            int x = s % 8, y = s / 8, nx, ny;
            f(x, y, nx, ny);
            out |= 1ULL << sq(nx, ny);
        }
    return out;
}

inline int popcnt(uint64_t b) { return __builtin_popcountll(b); }

uint32_t pow3(int n)
{
    uint32_t r = 1;
    for (int i = 0; i < n; i++) r *= 3;
    return r;
}

struct PO { uint64_t P, O; };

PO fV(PO p) { return {or_flip_vertical(p.P), or_flip_vertical(p.O)}; }
PO fH(PO p) { return {or_flip_horizontal(p.P), or_flip_horizontal(p.O)}; }
PO fD(PO p) { return {or_flip_diagonal(p.P), or_flip_diagonal(p.O)}; }
PO fC(PO p) { return {or_flip_codiagonal(p.P), or_flip_codiagonal(p.O)}; }

uint32_t fi(PO p, uint64_t m) { return or_fast_index(p.P, p.O, m); }

bool is_v(uint64_t m) { return or_flip_vertical(m) == m; }
bool is_h(uint64_t m) { return or_flip_horizontal(m) == m; }
bool is_d(uint64_t m) { return or_flip_diagonal(m) == m; }
bool is_c(uint64_t m) { return or_flip_codiagonal(m) == m; }

// ------------------------------------------------------------ indexers ----
// Pattern/Indexer.cpp:18-317.  One struct covers the six classes.
enum Cls { A = 0, V = 1, D = 2, VH = 3, DC = 4, VHDC = 5 };

constexpr uint64_t kHalfV = 0x00000000FFFFFFFFULL;  // Indexer.cpp:64 (rows 0-3)
constexpr uint64_t kHalfD = 0x0080C0E0F0F8FCFEULL;  // Indexer.cpp:107 (x > y)
constexpr uint64_t kDiag  = 0x8040201008040201ULL;  // Indexer.cpp:108 (x == y)

struct Pat {
    Cls cls;
    uint64_t mask;       // transformed mask kept by the indexer
    int order;           // symmetry_group_order
    int64_t size;        // index_space_size
    uint32_t hs = 0, ds = 0;
    std::vector<uint32_t> dense;  // VH / DC / VHDC

    uint32_t pair(uint32_t a, uint32_t b) const
    {
        uint32_t lo = std::min(a, b), hi = std::max(a, b);
        return lo * hs + hi - lo * (lo + 1) / 2;
    }

    uint32_t index(PO p) const
    {
        switch (cls) {
        case A: return fi(p, mask);
        case V: return pair(fi(p, mask & kHalfV), fi(fV(p), mask & kHalfV));
        case D: return pair(fi(p, mask & kHalfD), fi(fD(p), mask & kHalfD)) * ds + fi(p, mask & kDiag);
        default: return dense[fi(p, mask)];
        }
    }

    // InsertIndices: the symmetry applied per slot, in slot order.
    void indices(PO p, uint32_t* out, uint32_t off) const
    {
        switch (cls) {
        case A:  // Indexer.cpp:31-41
            out[0] = off + index(p);
            out[1] = off + index(fC(p));
            out[2] = off + index(fD(p));
            out[3] = off + index(fH(p));
            out[4] = off + index(fV(p));
            out[5] = off + index(fH(fC(p)));
            out[6] = off + index(fH(fD(p)));
            out[7] = off + index(fH(fV(p)));
            break;
        case V:  // Indexer.cpp:82-88
            out[0] = off + index(p);
            out[1] = off + index(fC(p));
            out[2] = off + index(fD(p));
            out[3] = off + index(fH(p));
            break;
        case D:  // Indexer.cpp:128-134
            out[0] = off + index(p);
            out[1] = off + index(fC(p));
            out[2] = off + index(fH(p));
            out[3] = off + index(fV(p));
            break;
        case VH:  // Indexer.cpp:186-190
            out[0] = off + index(p);
            out[1] = off + index(fD(p));
            break;
        case DC:  // Indexer.cpp:240-244
            out[0] = off + index(p);
            out[1] = off + index(fV(p));
            break;
        case VHDC:  // Indexer.cpp:306-309
            out[0] = off + index(p);
            break;
        }
    }

    std::vector<uint64_t> variations() const
    {
        uint64_t p = mask;
        auto C = or_flip_codiagonal; auto Dg = or_flip_diagonal;
        auto H = or_flip_horizontal; auto Vv = or_flip_vertical;
        switch (cls) {
        case A: return {p, C(p), Dg(p), H(p), Vv(p), C(H(p)), Dg(H(p)), Vv(H(p))};
        case V: return {p, C(p), Dg(p), H(p)};
        case D: return {p, C(p), H(p), Vv(p)};
        case VH: return {p, Dg(p)};
        case DC: return {p, Vv(p)};
        default: return {p};
        }
    }
};

// Dense numbering of the orbits for VH / DC / VHDC (Indexer.cpp:159-180,
// 213-234, 271-300): first-seen order of the Configurations enumeration.
void build_dense(Pat& pt)
{
    const uint64_t m = pt.mask;
    const int64_t n3 = pow3(popcnt(m));
    pt.dense.assign(n3, 0xFFFFFFFFu);
    int64_t cap = n3;
    std::vector<uint64_t> cfg(2 * cap);
    int64_t nc = or_configurations(m, cfg.data(), cap);
    uint32_t next = 0;
    for (int64_t i = 0; i < nc; i++) {
        PO c{cfg[2 * i], cfg[2 * i + 1]};
        if (pt.dense[fi(c, m)] != 0xFFFFFFFFu) continue;
        std::vector<PO> imgs;
        if (pt.cls == VH) imgs = {c, fV(c), fH(c), fH(fV(c))};
        else if (pt.cls == DC) imgs = {c, fD(c), fC(c), fC(fD(c))};
        else imgs = {c, fC(c), fD(c), fH(c), fV(c), fC(fH(c)), fD(fH(c)), fV(fH(c))};
        for (auto& im : imgs) pt.dense[fi(im, m)] = next;
        next++;
    }
    pt.size = next;
}

// CreateIndexer (Indexer.cpp:319-365)
Pat make_pattern(uint64_t p)
{
    bool v = is_v(p), h = is_h(p), d = is_d(p), c = is_c(p);
    Pat pt;
    if (v && h && d && c) { pt.cls = VHDC; pt.mask = p; pt.order = 1; }
    else if (v && h) { pt.cls = VH; pt.mask = p; pt.order = 2; }
    else if (d && c) { pt.cls = DC; pt.mask = p; pt.order = 2; }
    else if (v) { pt.cls = V; pt.mask = p; pt.order = 4; }
    else if (h) { pt.cls = V; pt.mask = or_flip_diagonal(p); pt.order = 4; }
    else if (d) { pt.cls = D; pt.mask = p; pt.order = 4; }
    else if (c) { pt.cls = D; pt.mask = or_flip_vertical(p); pt.order = 4; }
    else { pt.cls = A; pt.mask = p; pt.order = 8; }

    switch (pt.cls) {
    case A: pt.size = pow3(popcnt(pt.mask)); break;
    case V:
        pt.hs = pow3(popcnt(pt.mask & kHalfV));
        pt.size = int64_t(pt.hs) * (pt.hs + 1) / 2;
        break;
    case D:
        pt.hs = pow3(popcnt(pt.mask & kHalfD));
        pt.ds = pow3(popcnt(pt.mask & kDiag));
        pt.size = int64_t(pt.ds) * pt.hs * (pt.hs + 1) / 2;
        break;
    default: build_dense(pt); break;
    }
    return pt;
}

}  // namespace

struct or_group {
    std::vector<Pat> pats;
    std::vector<uint32_t> offsets;
    int k = 0;
    int64_t cols = 0;
};

extern "C" {

// ------------------------------------------------------------ bitboards ---
uint64_t or_flip_vertical(uint64_t b)
{
    return map_squares(b, [](int x, int y, int& nx, int& ny) { nx = x; ny = 7 - y; });
}
uint64_t or_flip_horizontal(uint64_t b)
{
    return map_squares(b, [](int x, int y, int& nx, int& ny) { nx = 7 - x; ny = y; });
}
uint64_t or_flip_diagonal(uint64_t b)
{
    return map_squares(b, [](int x, int y, int& nx, int& ny) { nx = y; ny = x; });
}
uint64_t or_flip_codiagonal(uint64_t b)
{
    return map_squares(b, [](int x, int y, int& nx, int& ny) { nx = 7 - y; ny = 7 - x; });
}

uint64_t or_pext(uint64_t src, uint64_t mask)
{
    uint64_t out = 0;
    int j = 0;
    for (int s = 0; s < 64; s++)
        if ((mask >> s) & 1) {
            out |= ((src >> s) & 1ULL) << j;
            j++;
        }
    return out;
}

uint64_t or_pdep(uint64_t src, uint64_t mask)
{
    uint64_t out = 0;
    int j = 0;
    for (int s = 0; s < 64; s++)
        if ((mask >> s) & 1) {
            out |= ((src >> j) & 1ULL) << s;
            j++;
        }
    return out;
}

uint32_t or_fast_index(uint64_t P, uint64_t O, uint64_t mask)
{
    uint32_t idx = 0, w = 1;
    for (int s = 0; s < 64; s++)
        if ((mask >> s) & 1) {
            idx += w * (uint32_t((P >> s) & 1) + 2u * uint32_t((O >> s) & 1));
            w *= 3;
        }
    return idx;
}

int64_t or_configurations(uint64_t mask, uint64_t* out, int64_t cap)
{
    const uint64_t n = 1ULL << popcnt(mask);
    int64_t c = 0;
    for (uint64_t p = 0; p < n; p++)
        for (uint64_t o = 0; o < n; o++) {
            if (p & o) continue;
            if (c < cap) {
                out[2 * c] = or_pdep(p, mask);
                out[2 * c + 1] = or_pdep(o, mask);
            }
            c++;
        }
    return c;
}

// ------------------------------------------------------------ group -------
or_group* or_group_create(const uint64_t* masks, int n)
{
    auto* g = new or_group;
    uint32_t off = 0;
    for (int i = 0; i < n; i++) {
        g->pats.push_back(make_pattern(masks[i]));
        g->offsets.push_back(off);
        off += uint32_t(g->pats.back().size);
        g->k += g->pats.back().order;
    }
    g->cols = off;
    return g;
}

void or_group_free(or_group* g) { delete g; }
int or_group_k(const or_group* g) { return g->k; }
int64_t or_group_cols(const or_group* g) { return g->cols; }
int or_group_npatterns(const or_group* g) { return int(g->pats.size()); }

void or_group_pattern_info(const or_group* g, int i, int* cls, uint64_t* mask,
                           int64_t* size, int* order)
{
    const Pat& p = g->pats[i];
    if (cls) *cls = p.cls;
    if (mask) *mask = p.mask;
    if (size) *size = p.size;
    if (order) *order = p.order;
}

void or_group_variations(const or_group* g, uint64_t* out)
{
    int j = 0;
    for (auto& p : g->pats)
        for (uint64_t v : p.variations()) out[j++] = v;
}

void or_group_indices(const or_group* g, const uint64_t* PO_, int64_t n, uint32_t* out)
{
    const int k = g->k;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) {
        PO p{PO_[2 * r], PO_[2 * r + 1]};
        uint32_t* row = out + r * k;
        int j = 0;
        for (size_t i = 0; i < g->pats.size(); i++) {
            g->pats[i].indices(p, row + j, g->offsets[i]);
            j += g->pats[i].order;
        }
    }
}

uint32_t or_group_pattern_index(const or_group* g, int i, uint64_t P, uint64_t O)
{
    return g->pats[i].index(PO{P, O});
}

// ------------------------------------------------------------ CSR ---------
void or_spmv_f32(const uint32_t* idx, int64_t rows, int k, const float* x, float* y)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; r++) {
        float s = 0.0f;
        for (int j = 0; j < k; j++) s += x[idx[r * k + j]];
        y[r] = s;
    }
}

void or_spmv_f64(const uint32_t* idx, int64_t rows, int k, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; r++) {
        double s = 0.0;
        for (int j = 0; j < k; j++) s += x[idx[r * k + j]];
        y[r] = s;
    }
}

void or_spmtv_f32(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                  const float* y, float* z, int nthreads)
{
    std::fill(z, z + cols, 0.0f);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        std::vector<float> local(cols, 0.0f);
#pragma omp for nowait schedule(static)
        for (int64_t r = 0; r < rows; r++)
            for (int j = 0; j < k; j++) local[idx[r * k + j]] += y[r];
#pragma omp critical
        for (int64_t c = 0; c < cols; c++) z[c] += local[c];
    }
}

void or_spmtv_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                  const double* y, double* z)
{
    // Private per-thread buffers reduced in thread order: deterministic for a
    // fixed thread count.
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<double>> part(nt);
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        std::vector<double>& local = part[t];
        local.assign(cols, 0.0);
        int64_t chunk = (rows + nt - 1) / nt;
        int64_t lo = std::min<int64_t>(rows, t * chunk), hi = std::min<int64_t>(rows, lo + chunk);
        for (int64_t r = lo; r < hi; r++)
            for (int j = 0; j < k; j++) local[idx[r * k + j]] += y[r];
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < cols; c++) {
        double s = 0.0;
        for (int t = 0; t < nt; t++) s += part[t][c];
        z[c] = s;
    }
}

void or_jacobi_f32(const uint32_t* idx, int64_t rows, int k, int64_t cols, float* d)
{
    std::vector<float> diag(cols, 0.0f);
    std::vector<uint32_t> row(k);
    for (int64_t r = 0; r < rows; r++) {
        std::copy(idx + r * k, idx + r * k + k, row.begin());
        std::sort(row.begin(), row.end());
        for (int i = 0; i < k;) {
            int j = i + 1;
            while (j < k && row[j] == row[i]) j++;
            diag[row[i]] += float((j - i) * (j - i));
            i = j;
        }
    }
    for (int64_t c = 0; c < cols; c++) d[c] = 1.0f / diag[c];
}

}  // extern "C"

// ------------------------------------------------------------ solvers -----
namespace {

template <class T>
struct Csr {
    const uint32_t* idx;
    int64_t rows;
    int k;
    int64_t cols;

    void mul(const std::vector<T>& x, std::vector<T>& y) const;   // y = A x
    void tmul(const std::vector<T>& y, std::vector<T>& z) const;  // z = A^T y
};

template <>
void Csr<float>::mul(const std::vector<float>& x, std::vector<float>& y) const
{
    y.resize(rows);
    or_spmv_f32(idx, rows, k, x.data(), y.data());
}
template <>
void Csr<float>::tmul(const std::vector<float>& y, std::vector<float>& z) const
{
    z.resize(cols);
    or_spmtv_f32(idx, rows, k, cols, y.data(), z.data(), 0);
}
template <>
void Csr<double>::mul(const std::vector<double>& x, std::vector<double>& y) const
{
    y.resize(rows);
    or_spmv_f64(idx, rows, k, x.data(), y.data());
}
template <>
void Csr<double>::tmul(const std::vector<double>& y, std::vector<double>& z) const
{
    z.resize(cols);
    or_spmtv_f64(idx, rows, k, cols, y.data(), z.data());
}

// Vector helpers with float reductions for T=float (Math/Vector.cpp:126-142).
template <class T>
T vdot(const std::vector<T>& a, const std::vector<T>& b)
{
    T s = 0;
    const int64_t n = int64_t(a.size());
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

template <class T>
T vnorm(const std::vector<T>& a) { return std::sqrt(vdot(a, a)); }

// x += a*y
template <class T>
void axpy(std::vector<T>& x, T a, const std::vector<T>& y)
{
    const int64_t n = int64_t(x.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) x[i] += a * y[i];
}

// Jacobi diagonal with non-normal values replaced by 1 (Solver.h:25-38).
template <class T>
std::vector<T> jacobi(const Csr<T>& A)
{
    std::vector<float> d(A.cols);
    or_jacobi_f32(A.idx, A.rows, A.k, A.cols, d.data());
    std::vector<T> out(A.cols);
    for (int64_t c = 0; c < A.cols; c++) out[c] = std::isnormal(d[c]) ? T(d[c]) : T(1);
    return out;
}

// PCG on A^T A (Solver.h:78-117 with TwoMatrix, MatrixProxy.h:63-67).
template <class T>
void pcg(const Csr<T>& A, const T* b, int iters, T* xout, T* hist)
{
    std::vector<T> P = jacobi(A);
    std::vector<T> bb(b, b + A.rows), rhs, x(A.cols, T(0)), r, z(A.cols), p, Ap, tmp;
    A.tmul(bb, rhs);            // transposed(A) * b
    r = rhs;                    // r = rhs - AtA*0
    for (int64_t c = 0; c < A.cols; c++) z[c] = P[c] * r[c];
    p = z;
    for (int it = 0; it < iters; it++) {
        T rz_old = vdot(r, z);
        A.mul(p, tmp);
        A.tmul(tmp, Ap);
        T alpha = rz_old / vdot(p, Ap);
        axpy(x, alpha, p);
        axpy(r, -alpha, Ap);
        for (int64_t c = 0; c < A.cols; c++) z[c] = P[c] * r[c];
        T beta = vdot(r, z) / rz_old;
        for (int64_t c = 0; c < A.cols; c++) p[c] = z[c] + beta * p[c];
        if (hist) hist[it] = vdot(r, r);
    }
    std::copy(x.begin(), x.end(), xout);
}

template <class T>
void cg(const Csr<T>& A, const T* b, int iters, T* xout, T* hist)
{
    std::vector<T> bb(b, b + A.rows), r, x(A.cols, T(0)), p, Ap, tmp;
    A.tmul(bb, r);
    p = r;
    for (int it = 0; it < iters; it++) {
        T rr_old = vdot(r, r);
        A.mul(p, tmp);
        A.tmul(tmp, Ap);
        T alpha = rr_old / vdot(p, Ap);
        axpy(x, alpha, p);
        axpy(r, -alpha, Ap);
        T beta = vdot(r, r) / rr_old;
        for (int64_t c = 0; c < A.cols; c++) p[c] = r[c] + beta * p[c];
        if (hist) hist[it] = vdot(r, r);
    }
    std::copy(x.begin(), x.end(), xout);
}

// decompose (Solver.h:119-126): (|v|, v/|v|), or (0, 0) for v = 0.
template <class T>
T normalize(std::vector<T>& v)
{
    T n = vnorm(v);
    if (n == T(0)) {
        std::fill(v.begin(), v.end(), T(0));
        return T(0);
    }
    for (auto& e : v) e = e / n;
    return n;
}

// LSQR (Solver.h:128-182), optional column scaling D: works on A D.
template <class T>
void lsqr(const Csr<T>& A, const T* b, int iters, const T* D, T* xout, T* phist, T* nres)
{
    std::vector<T> u(b, b + A.rows), v, w, x(A.cols, T(0)), t, s(A.cols);
    T beta = normalize(u);
    A.tmul(u, v);
    if (D) for (int64_t c = 0; c < A.cols; c++) v[c] *= D[c];
    T alpha = normalize(v);
    w = v;
    T phibar = beta, rhobar = alpha;
    for (int it = 0; it < iters; it++) {
        // u = A v - alpha u
        if (D) {
            for (int64_t c = 0; c < A.cols; c++) s[c] = D[c] * v[c];
            A.mul(s, t);
        } else {
            A.mul(v, t);
        }
        for (int64_t r = 0; r < A.rows; r++) t[r] = t[r] - alpha * u[r];
        u.swap(t);
        beta = normalize(u);
        // v = A^T u - beta v
        A.tmul(u, t);
        if (D) for (int64_t c = 0; c < A.cols; c++) t[c] *= D[c];
        for (int64_t c = 0; c < A.cols; c++) t[c] = t[c] - beta * v[c];
        v.swap(t);
        alpha = normalize(v);

        T rho = std::sqrt(rhobar * rhobar + beta * beta);
        T c = rhobar / rho;
        T sn = beta / rho;
        T theta = sn * alpha;
        rhobar = -c * alpha;
        T phi = c * phibar;
        phibar = sn * phibar;
        axpy(x, phi / rho, w);
        for (int64_t j = 0; j < A.cols; j++) w[j] = v[j] - (theta / rho) * w[j];
        if (phist) phist[it] = phibar;
        if (nres) nres[it] = phibar * alpha * std::abs(c);
    }
    for (int64_t c = 0; c < A.cols; c++) xout[c] = D ? D[c] * x[c] : x[c];
}

// Symmetric Givens rotation (the form Solver.h:193-217 takes from scipy).
template <class T>
void sym_ortho(T a, T b, T& c, T& s, T& r)
{
    auto sgn = [](T v) { return v < 0 ? T(-1) : (v > 0 ? T(1) : T(0)); };
    if (b == 0) { c = sgn(a); s = 0; r = std::abs(a); return; }
    if (a == 0) { c = 0; s = sgn(b); r = std::abs(b); return; }
    if (std::abs(b) > std::abs(a)) {
        T tau = a / b;
        s = sgn(b) / std::sqrt(1 + tau * tau);
        c = s * tau;
        r = b / s;
    } else {
        T tau = b / a;
        c = sgn(a) / std::sqrt(1 + tau * tau);
        s = c * tau;
        r = a / c;
    }
}

// LSMR (Fong & Saunders 2011, damp = 0), following Solver.h:219-287.
template <class T>
void lsmr(const Csr<T>& A, const T* b, int iters, T* xout)
{
    std::vector<T> u(b, b + A.rows), v, x(A.cols, T(0)), t;
    T beta = normalize(u);
    A.tmul(u, v);
    T alpha = normalize(v);
    T alphabar = alpha, zetabar = alpha * beta, rho = 1, rhobar = 1, cbar = 1, sbar = 0;
    std::vector<T> h = v, hbar(A.cols, T(0));
    for (int it = 0; it < iters; it++) {
        T rho_old = rho, rhobar_old = rhobar;
        A.mul(v, t);
        for (int64_t r = 0; r < A.rows; r++) t[r] -= alpha * u[r];
        u.swap(t);
        beta = normalize(u);
        A.tmul(u, t);
        for (int64_t c = 0; c < A.cols; c++) t[c] -= beta * v[c];
        v.swap(t);
        alpha = normalize(v);

        T c, s;
        sym_ortho(alphabar, beta, c, s, rho);
        T theta_new = s * alpha;
        alphabar = c * alpha;
        T thetabar = sbar * rho;
        sym_ortho(cbar * rho, theta_new, cbar, sbar, rhobar);
        T zeta = cbar * zetabar;
        zetabar = -sbar * zetabar;
        T f1 = thetabar * rho / (rho_old * rhobar_old);
        for (int64_t j = 0; j < A.cols; j++) hbar[j] = h[j] - f1 * hbar[j];
        axpy(x, zeta / (rho * rhobar), hbar);
        T f2 = theta_new / rho;
        for (int64_t j = 0; j < A.cols; j++) h[j] = v[j] - f2 * h[j];
    }
    std::copy(x.begin(), x.end(), xout);
}

}  // namespace

extern "C" {

void or_pcg_f32(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                const float* b, int iterations, float* x, float* hist)
{
    pcg(Csr<float>{idx, rows, k, cols}, b, iterations, x, hist);
}
void or_pcg_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                const double* b, int iterations, double* x, double* hist)
{
    pcg(Csr<double>{idx, rows, k, cols}, b, iterations, x, hist);
}
void or_cg_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
               const double* b, int iterations, double* x, double* hist)
{
    cg(Csr<double>{idx, rows, k, cols}, b, iterations, x, hist);
}
void or_lsqr_f32(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                 const float* b, int iterations, float* x, float* phist)
{
    lsqr<float>(Csr<float>{idx, rows, k, cols}, b, iterations, nullptr, x, phist, nullptr);
}
void or_lsqr_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                 const double* b, int iterations, const double* colscale,
                 double* x, double* phist, double* nres)
{
    lsqr<double>(Csr<double>{idx, rows, k, cols}, b, iterations, colscale, x, phist, nres);
}
void or_lsmr_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                 const double* b, int iterations, double* x)
{
    lsmr(Csr<double>{idx, rows, k, cols}, b, iterations, x);
}

double or_rel_normal_residual(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                              const double* b, const double* x,
                              double* res_norm, double* atb_norm)
{
    std::vector<double> ax(rows), r(rows), g(cols), atb(cols);
    or_spmv_f64(idx, rows, k, x, ax.data());
    for (int64_t i = 0; i < rows; i++) r[i] = b[i] - ax[i];
    or_spmtv_f64(idx, rows, k, cols, r.data(), g.data());
    or_spmtv_f64(idx, rows, k, cols, b, atb.data());
    double gn = 0, an = 0, rn = 0;
    for (int64_t c = 0; c < cols; c++) { gn += g[c] * g[c]; an += atb[c] * atb[c]; }
    for (int64_t i = 0; i < rows; i++) rn += r[i] * r[i];
    if (res_norm) *res_norm = std::sqrt(rn);
    if (atb_norm) *atb_norm = std::sqrt(an);
    return an > 0 ? std::sqrt(gn) / std::sqrt(an) : std::sqrt(gn);
}

// ------------------------------------------------------------ input side --
// Loop-based rays, the form BoardTest/Flips.cpp:6-36 uses as its reference.
uint64_t or_flips(uint64_t P, uint64_t O, int move)
{
    uint64_t flips = 0;
    const int mx = move % 8, my = move / 8;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            if (!dx && !dy) continue;
            uint64_t run = 0;
            int x = mx + dx, y = my + dy;
            while (x >= 0 && x < 8 && y >= 0 && y < 8) {
                uint64_t bit = 1ULL << sq(x, y);
                if (O & bit) run |= bit;
                else {
                    if (P & bit) flips |= run;
                    break;
                }
                x += dx;
                y += dy;
            }
        }
    return flips;
}

uint64_t or_possible_moves(uint64_t P, uint64_t O)
{
    uint64_t empty = ~(P | O), moves = 0;
    for (int s = 0; s < 64; s++)
        if (((empty >> s) & 1) && or_flips(P, O, s)) moves |= 1ULL << s;
    return moves;
}

}  // extern "C"

namespace {

PO start_position() { return {0x0000000810000000ULL, 0x0000001008000000ULL}; }

PO play(PO p, int move)
{
    uint64_t f = or_flips(p.P, p.O, move);
    return {p.O ^ f, p.P ^ f ^ (1ULL << move)};
}

// Game/GameClass.cpp:4-13
PO implicit_pass(PO p)
{
    if (!or_possible_moves(p.P, p.O)) {
        PO q{p.O, p.P};
        if (or_possible_moves(q.P, q.O)) return q;
    }
    return p;
}

uint64_t perft(PO p, int plies)
{
    if (plies == 0) return 1;
    uint64_t m = or_possible_moves(p.P, p.O);
    uint64_t n = 0;
    for (int s = 0; s < 64; s++)
        if ((m >> s) & 1) n += perft(implicit_pass(play(p, s)), plies - 1);
    return n;
}

struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    // floor(next * n / 2^64)
    uint32_t below(uint32_t n) { return uint32_t(((unsigned __int128)next() * n) >> 64); }
    // Irwin-Hall: sum of 12 uniforms on [0,1) with 32-bit resolution, minus 6
    double ih12()
    {
        double s = 0;
        for (int i = 0; i < 12; i++) s += double(uint32_t(next() >> 32)) * (1.0 / 4294967296.0);
        return s - 6.0;
    }
};

uint64_t stream_seed(uint64_t seed, uint64_t i) { return seed ^ (0xD1B54A32D192ED03ULL * (i + 1)); }

}  // namespace

extern "C" {

uint64_t or_perft(int plies) { return perft(start_position(), plies); }

void or_random_positions_mt(uint32_t seed, int64_t n, uint64_t* out)
{
    std::mt19937_64 eng(seed);
    for (int64_t i = 0; i < n; i++) {
        uint64_t a = eng(), b = eng();
        out[2 * i] = a & ~b;
        out[2 * i + 1] = ~a & b;
    }
}

void or_random_positions_mt_dist(uint32_t seed, int64_t n, uint64_t* out)
{
    // Board/PositionGenerator.h:9-10 verbatim in behaviour: the draws go
    // through uniform_int_distribution<uint64_t>{0, 2^64-1}.
    std::mt19937_64 eng(seed);
    std::uniform_int_distribution<uint64_t> dist{0, 0xFFFFFFFFFFFFFFFFULL};
    for (int64_t i = 0; i < n; i++) {
        uint64_t a = dist(eng), b = dist(eng);
        out[2 * i] = a & ~b;
        out[2 * i + 1] = ~a & b;
    }
}

void or_random_play(uint64_t seed, int64_t n, int e_lo, int e_hi, uint64_t* out)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        SplitMix rng{stream_seed(seed, uint64_t(i))};
        int E = e_lo == e_hi ? e_lo : e_lo + int(rng.below(uint32_t(e_hi - e_lo + 1)));
        PO p = start_position();
        for (;;) {
            if (popcnt(~(p.P | p.O)) == E) break;
            uint64_t m = or_possible_moves(p.P, p.O);
            if (!m) { p = start_position(); continue; }
            uint32_t r = rng.below(uint32_t(popcnt(m)));
            int mv = -1;
            for (int s = 0, seen = 0; s < 64; s++)
                if ((m >> s) & 1) {
                    if (seen == int(r)) { mv = s; break; }
                    seen++;
                }
            p = implicit_pass(play(p, mv));
        }
        out[2 * i] = p.P;
        out[2 * i + 1] = p.O;
    }
}

void or_make_targets(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                     uint64_t seed_x, uint64_t seed_n, double sx, double sn,
                     double* x_true, double* b)
{
    for (int64_t c = 0; c < cols; c++) {
        SplitMix rng{stream_seed(seed_x, uint64_t(c))};
        x_true[c] = sx * rng.ih12();
    }
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; r++) {
        double s = 0.0;
        for (int j = 0; j < k; j++) s += x_true[idx[r * k + j]];
        SplitMix rng{stream_seed(seed_n, uint64_t(r))};
        double v = s + sn * rng.ih12();
        double e = 2.0 * std::floor(v * 0.5 + 0.5);
        b[r] = std::min(64.0, std::max(-64.0, e));
    }
}

// ------------------------------------------------------------ evaluator ---
int64_t or_expand_weights(const or_group* g, const float* w, float* out)
{
    int64_t total = 0;
    for (size_t i = 0; i < g->pats.size(); i++) {
        const Pat& p = g->pats[i];
        auto vars = p.variations();
        const int64_t n3 = pow3(popcnt(p.mask));
        for (size_t v = 0; v < vars.size(); v++) {
            if (out) {
                float* tab = out + total;
                std::fill(tab, tab + n3, 0.0f);
                std::vector<uint64_t> cfg(2 * n3);
                int64_t nc = or_configurations(vars[v], cfg.data(), n3);
                std::vector<uint32_t> ids(p.order);
                for (int64_t c = 0; c < nc; c++) {
                    PO q{cfg[2 * c], cfg[2 * c + 1]};
                    p.indices(q, ids.data(), 0);
                    tab[fi(q, vars[v])] = w[g->offsets[i] + ids[v]];
                }
            }
            total += n3;
        }
    }
    return total;
}

void or_eval_expanded(const or_group* g, const float* tables, const uint64_t* PO_,
                      int64_t n, float* out)
{
    std::vector<uint64_t> masks;
    std::vector<int64_t> offs;
    int64_t off = 0;
    for (auto& p : g->pats)
        for (uint64_t v : p.variations()) {
            masks.push_back(v);
            offs.push_back(off);
            off += pow3(popcnt(v));
        }
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) {
        PO q{PO_[2 * r], PO_[2 * r + 1]};
        float s = 0.0f;
        for (size_t i = 0; i < masks.size(); i++) s += tables[offs[i] + fi(q, masks[i])];
        out[r] = s;
    }
}

void or_compact_weights(const or_group* g, const float* tables, float* w)
{
    int64_t off = 0;
    for (size_t i = 0; i < g->pats.size(); i++) {
        const Pat& p = g->pats[i];
        const int64_t n3 = pow3(popcnt(p.mask));
        std::vector<uint64_t> cfg(2 * n3);
        int64_t nc = or_configurations(p.mask, cfg.data(), n3);
        for (int64_t c = 0; c < p.size; c++) w[g->offsets[i] + c] = 0.0f;
        for (int64_t c = 0; c < nc; c++) {
            PO q{cfg[2 * c], cfg[2 * c + 1]};
            w[g->offsets[i] + p.index(q)] = tables[off + fi(q, p.mask)];
        }
        off += n3 * int64_t(p.variations().size());
    }
}

}  // extern "C"
