// FP64 CPU oracle for the WCSPH step -- TEST INFRASTRUCTURE, never shipped.
//
// A restatement (own code) of the RRSPH_OMP step of RackotRR/SPH2D.  Each
// function cites the reference file:line whose behaviour it reproduces; the
// arithmetic is kept in the same operation order as the reference so that the
// FP64 results agree to the last few ulps.
#include "sph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint64_t kInvalidCell = 0xFFFFFFFFull;  // GridUtils.h:8 (UINT_MAX)
constexpr uint64_t kMaxCellCoord2 = 65535;        // GridUtils.h:9

template <int D> struct V;
template <> struct V<2> {
    double x = 0, y = 0;
    V() = default;
    V(double a, double b) : x(a), y(b) {}
    static V load(const double* p, size_t i) { return V(p[2 * i], p[2 * i + 1]); }
    void store(double* p, size_t i) const { p[2 * i] = x; p[2 * i + 1] = y; }
};
template <> struct V<3> {
    double x = 0, y = 0, z = 0;
    V() = default;
    V(double a, double b, double c) : x(a), y(b), z(c) {}
    static V load(const double* p, size_t i) { return V(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
    void store(double* p, size_t i) const { p[3 * i] = x; p[3 * i + 1] = y; p[3 * i + 2] = z; }
};
inline V<2> operator+(V<2> a, V<2> b) { return {a.x + b.x, a.y + b.y}; }
inline V<2> operator-(V<2> a, V<2> b) { return {a.x - b.x, a.y - b.y}; }
inline V<2> operator-(V<2> a) { return {-a.x, -a.y}; }
inline V<2> operator*(V<2> a, double s) { return {a.x * s, a.y * s}; }
inline V<2> operator/(V<2> a, double s) { return {a.x / s, a.y / s}; }
inline double dot(V<2> a, V<2> b) { return a.x * b.x + a.y * b.y; }
inline V<3> operator+(V<3> a, V<3> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V<3> operator-(V<3> a, V<3> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V<3> operator-(V<3> a) { return {-a.x, -a.y, -a.z}; }
inline V<3> operator*(V<3> a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V<3> operator/(V<3> a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V<3> a, V<3> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <int D> inline V<D>& operator+=(V<D>& a, V<D> b) { a = a + b; return a; }
template <int D> inline double len2(V<D> a) { return dot(a, a); }
template <int D> inline double len(V<D> a) { return std::sqrt(len2(a)); }
template <int D> inline V<D> zero() { return V<D>(); }

inline double sq(double v) { return v * v; }
inline double cb(double v) { return v * v * v; }
// repeated multiplication, CommonIncl.h:36-42
inline double mulpow(double v, unsigned n) {
    double r = 1;
    for (unsigned k = n; k > 0; --k) r *= v;
    return r;
}

// ---------------------------------------------------------------- kernels
// SmoothingKernel.h:14-283 (factors, W and grad W w.r.t. diff = r_i - r_j)
struct Kern {
    const oracle_params& P;
    double factor(int skf) const {
        const double h = P.hsml, pi = P.pi;
        const bool d3 = P.dim == 3;
        switch (skf) {
        case 2: return d3 ? 1 / (pi * std::sqrt(pi) * cb(h)) : 1 / (pi * sq(h));
        case 3: return d3 ? 21 / (16 * pi * cb(h)) : 7 / (4 * pi * sq(h));
        case 4: return d3 ? 15 / (pi * cb(4 * h)) : 5 / (16 * pi * sq(h));
        default: return d3 ? double(3) / (double(2) * pi * cb(h)) : double(15) / (double(7) * pi * sq(h));
        }
    }
    double w(double dist, int skf) const {
        const double q = dist / P.hsml;
        const double f = factor(skf);
        switch (skf) {
        case 2: return f * std::exp(-q * q);
        case 3: return q <= 2 ? f * mulpow(1 - 0.5 * q, 4) * (2 * q + 1) : 0;
        case 4: return q <= 2 ? f * cb(2 - q) : 0;
        default:
            if (q <= 1) return f * (double(2) / double(3) - sq(q) + cb(q) * 0.5);
            if (q <= 2) return f * (double(1) / double(6) * cb(2 - q));
            return 0;
        }
    }
    template <int D> V<D> dwdr(double dist, V<D> diff, int skf) const {
        const double h = P.hsml;
        const double q = dist / h;
        const double f = factor(skf);
        switch (skf) {
        case 2: return diff / sq(h) * (-2) * w(dist, 2);
        case 3:
            if (q < 1.e-6f) return zero<D>();
            if (q < 2) {
                V<D> g = diff / h / dist * f * 2;
                return g * (mulpow(1 - 0.5 * q, 4) - (2 * q + 1) * mulpow(1 - 0.5 * q, 3));
            }
            return zero<D>();
        case 4:
            if (q < 2) return diff / dist / h * f * (-3 * sq(2 - q));
            return zero<D>();
        default:
            if (q <= 1) return diff / sq(h) * f * (-double(2) + double(3) * 0.5 * q);
            if (q <= 2) return -diff / dist / h * f * (sq(2 - q) * 0.5);
            return zero<D>();
        }
    }
};

// ---------------------------------------------------------------- EOS
// EOS.cpp:3-44 -- Tait, gamma = 7, B = c^2 rho0 / 7
inline double eos_B(const oracle_params& P) { return P.eos_c_sqr * P.rho0 / 7; }
inline double eos_p(const oracle_params& P, double rho) {
    return eos_B(P) * (mulpow(rho / P.rho0, 7) - 1);
}

// ---------------------------------------------------------------- grid
// GridUtils.h:46-125 magic-bit Morton codes
inline uint32_t split2(uint32_t x) {
    x &= 0x0000FFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
inline uint32_t join2(uint32_t m) {
    uint32_t x = m & 0x55555555u;
    x = (x ^ (x >> 1)) & 0x33333333u;
    x = (x ^ (x >> 2)) & 0x0F0F0F0Fu;
    x = (x ^ (x >> 4)) & 0x00FF00FFu;
    x = (x ^ (x >> 8)) & 0x0000FFFFu;
    return x;
}
inline uint32_t split3(uint32_t a) {
    uint32_t x = a & 0x000003FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
inline uint32_t join3(uint32_t m) {
    uint32_t x = m & 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030C30C3u;
    x = (x ^ (x >> 4)) & 0x0300F00Fu;
    x = (x ^ (x >> 8)) & 0x030000FFu;
    x = (x ^ (x >> 16)) & 0x000003FFu;
    return x;
}

inline double cell_size(const oracle_params& P) { return P.cell_scale_k * P.hsml; }
// GridUtils.h:135-137: truncating conversion of (r - rmin)/cell; the
// reference's rr_uint is 64-bit on LP64, Morton then keeps the low bits.
inline uint64_t coord_of(const oracle_params& P, double r, double rmin) {
    return (uint64_t)((r - rmin) / cell_size(P));
}
template <int D> uint64_t cell_of(const oracle_params& P, V<D> r);
template <> uint64_t cell_of<2>(const oracle_params& P, V<2> r) {
    uint64_t cx = coord_of(P, r.x, P.x_mingeom), cy = coord_of(P, r.y, P.y_mingeom);
    return split2((uint32_t)cx) | (split2((uint32_t)cy) << 1);
}
template <> uint64_t cell_of<3>(const oracle_params& P, V<3> r) {
    uint64_t cx = coord_of(P, r.x, P.x_mingeom), cy = coord_of(P, r.y, P.y_mingeom),
             cz = coord_of(P, r.z, P.z_mingeom);
    return split3((uint32_t)cx) | (split3((uint32_t)cy) << 1) | (split3((uint32_t)cz) << 2);
}
inline uint64_t up(uint64_t c) { return c == kMaxCellCoord2 ? kInvalidCell : c + 1; }
inline uint64_t down(uint64_t c) { return c == 0 ? kInvalidCell : c - 1; }

// GridUtils.h:190-250: neighbour cells in x-major order
void neighbour_cells(int dim, uint64_t idx, uint64_t* cells) {
    const int n = dim == 3 ? 27 : 9;
    if (idx == kInvalidCell) {
        for (int k = 0; k < n; ++k) cells[k] = kInvalidCell;
        return;
    }
    if (dim == 2) {
        uint64_t cx = join2((uint32_t)idx), cy = join2((uint32_t)idx >> 1);
        uint64_t xs[3] = {down(cx), cx, up(cx)}, ys[3] = {down(cy), cy, up(cy)};
        int k = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                cells[k++] = (xs[a] == kInvalidCell || ys[b] == kInvalidCell)
                                 ? kInvalidCell
                                 : (uint64_t)(split2((uint32_t)xs[a]) | (split2((uint32_t)ys[b]) << 1));
    } else {
        uint64_t cx = join3((uint32_t)idx), cy = join3((uint32_t)idx >> 1), cz = join3((uint32_t)idx >> 2);
        uint64_t xs[3] = {down(cx), cx, up(cx)}, ys[3] = {down(cy), cy, up(cy)}, zs[3] = {down(cz), cz, up(cz)};
        int k = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c)
                    cells[k++] = (xs[a] == kInvalidCell || ys[b] == kInvalidCell || zs[c] == kInvalidCell)
                                     ? kInvalidCell
                                     : (uint64_t)(split3((uint32_t)xs[a]) | (split3((uint32_t)ys[b]) << 1) |
                                                  (split3((uint32_t)zs[c]) << 2));
    }
}

uint32_t bitlen_pow2(uint64_t id) {  // 1 << (intlog2(id) + 1), Input.cpp:72
    uint64_t passes = 0, s = id;
    while (s > 1) { s >>= 1; ++passes; }
    return (uint32_t)passes + 1;
}

// Input.cpp:45-75 (countCells) -- returns max_cells, or 0 with status 4
uint64_t count_cells(const oracle_params& P, int32_t* status) {
    uint64_t x_id = coord_of(P, P.x_maxgeom, P.x_mingeom);
    uint64_t y_id = coord_of(P, P.y_maxgeom, P.y_mingeom);
    uint64_t z_id = 0, max_id, id;
    if (P.dim == 3) {
        z_id = coord_of(P, P.z_maxgeom, P.z_mingeom);
        max_id = 1u << 10;
        id = split3((uint32_t)x_id) | (split3((uint32_t)y_id) << 1) | (split3((uint32_t)z_id) << 2);
    } else {
        max_id = 1u << 16;
        id = split2((uint32_t)x_id) | (split2((uint32_t)y_id) << 1);
    }
    if (x_id > max_id || y_id > max_id || z_id >= max_id) {
        if (status) *status = 4;
        return 0;
    }
    if (status) *status = 0;
    return 1ull << bitlen_pow2(id);
}

// Dense start table large enough for every neighbour cell of an in-domain
// particle: the reference reads cells_start(c + 1) for neighbour cells up to
// (x_id + 1, y_id + 1), which can lie past max_cells; those cells are empty.
uint64_t table_size(const oracle_params& P) {
    int32_t st = 0;
    uint64_t mc = count_cells(P, &st);
    uint64_t x_id = coord_of(P, P.x_maxgeom, P.x_mingeom) + 1;
    uint64_t y_id = coord_of(P, P.y_maxgeom, P.y_mingeom) + 1;
    uint64_t corner;
    if (P.dim == 3) {
        uint64_t z_id = coord_of(P, P.z_maxgeom, P.z_mingeom) + 1;
        corner = split3((uint32_t)std::min<uint64_t>(x_id, 1023)) |
                 (split3((uint32_t)std::min<uint64_t>(y_id, 1023)) << 1) |
                 (split3((uint32_t)std::min<uint64_t>(z_id, 1023)) << 2);
    } else {
        corner = split2((uint32_t)std::min<uint64_t>(x_id, 65535)) |
                 (split2((uint32_t)std::min<uint64_t>(y_id, 65535)) << 1);
    }
    return std::max<uint64_t>(mc, corner + 1) + 2;
}

template <int D>
int32_t grid_find_t(const oracle_params& P, const double* r, const int32_t* itype, uint32_t* nbr,
                    std::vector<uint32_t>& grid, std::vector<uint32_t>& start,
                    std::vector<uint64_t>& key) {
    const uint32_t n = P.ntotal;
    const uint64_t tsz = table_size(P);
    grid.assign(n, 0);
    key.assign(n, 0);
    start.assign(tsz, 0);
    // make_grid, GridFind.cpp:47-88: serial counting sort, stable scatter
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t c = cell_of<D>(P, V<D>::load(r, i));
        if (c >= tsz - 1) return 4;  // reference throws when consistency_check, else UB
        key[i] = c;
        start[c]++;
    }
    for (uint64_t c = 1; c < tsz; ++c) start[c] += start[c - 1];
    for (uint32_t k = n; k > 0; --k) {
        uint32_t j = k - 1;
        grid[--start[key[j]]] = j;
    }
    // find_neighbours, GridFind.cpp:112-189
    const int ncell = D == 3 ? 27 : 9;
    const double max_d2 = sq(cell_size(P));
    const uint32_t maxnb = P.max_neighbours;
    int32_t err = 0;
#pragma omp parallel for schedule(static) reduction(max : err)
    for (int64_t jj = 0; jj < (int64_t)n; ++jj) {
        const uint32_t j = (uint32_t)jj;
        uint32_t cnt = 0;
        if (itype[j] != 0) {
            uint64_t cells[27];
            V<D> rj = V<D>::load(r, j);
            neighbour_cells(D, key[j], cells);
            for (int k = 0; k < ncell; ++k) {
                uint64_t c = cells[k];
                if (c == kInvalidCell || c + 1 >= tsz) continue;
                for (uint32_t g = start[c]; g < start[c + 1]; ++g) {
                    uint32_t i = grid[g];
                    if (i == j || itype[i] == 0) continue;
                    V<D> d = V<D>::load(r, i) - rj;
                    if (len2(d) < max_d2) {
                        if (cnt == maxnb - 1) {  // overflow: the reference throws (consistency_check)
                            if (P.consistency_check) err = 1;
                            continue;  // without the check the extra neighbours are dropped
                        }
                        nbr[cnt + (size_t)maxnb * j] = i;
                        ++cnt;
                    }
                }
            }
        }
        uint32_t slot = std::min(cnt, maxnb - 1);
        nbr[slot + (size_t)maxnb * j] = n;
    }
    return err;
}

template <int D, class F>
inline void for_pairs(const oracle_params& P, const uint32_t* nbr, uint32_t j, F&& f) {
    const size_t base = (size_t)P.max_neighbours * j;
    for (uint32_t k = 0;; ++k) {
        uint32_t i = nbr[base + k];
        if (i == P.ntotal) break;
        f(i);
    }
}

// Density.h:25-61
template <int D>
// With rho_old (boundary-deficiency correction, docs/latex/ModelParams.tex:
// 124-133): rho_j = sum_i m W_ji / sum_i (m / rho_old_i) W_ji, both sums over
// j itself and its neighbours.  rho may not alias rho_old.
void density_sum_t(const oracle_params& P, const double* r, const uint32_t* nbr, double* rho, double* pres,
                   const double* rho_old = nullptr) {
    Kern K{P};
#pragma omp parallel for schedule(static)
    for (int64_t jj = 0; jj < (int64_t)P.ntotal; ++jj) {
        const uint32_t j = (uint32_t)jj;
        const double w0 = K.w(0, P.density_skf);
        double s = P.mass * w0;
        double sw = rho_old ? P.mass / rho_old[j] * w0 : 0.0;
        V<D> rj = V<D>::load(r, j);
        for_pairs<D>(P, nbr, j, [&](uint32_t i) {
            V<D> d = V<D>::load(r, i) - rj;
            const double w = K.w(len(d), P.density_skf);
            s += P.mass * w;
            if (rho_old) sw += P.mass / rho_old[i] * w;
        });
        if (rho_old) s = s / sw;
        rho[j] = s;
        pres[j] = eos_p(P, s);
    }
}

// Density.h:63-114
template <int D>
void density_con_t(const oracle_params& P, const double* r, const double* v, const uint32_t* nbr,
                   const double* rho, double* drho, double* pres) {
    Kern K{P};
    const double dhc = P.density_delta_sph_coef * P.hsml * P.eos_c;
#pragma omp parallel for schedule(static)
    for (int64_t jj = 0; jj < (int64_t)P.ntotal; ++jj) {
        const uint32_t j = (uint32_t)jj;
        double s = 0;
        V<D> rj = V<D>::load(r, j), vj = V<D>::load(v, j);
        for_pairs<D>(P, nbr, j, [&](uint32_t i) {
            V<D> ri = V<D>::load(r, i);
            V<D> diff = ri - rj;
            V<D> g = K.dwdr<D>(len(diff), diff, P.density_skf);
            V<D> dv = V<D>::load(v, i) - vj;
            s += P.mass * dot(dv, g);
            if (P.density_treatment == 2) {
                V<D> rab = rj - ri;
                double rf = dot(rab, g) / len2(rab);
                double pf = (rho[i] - rho[j]) / rho[i];
                s += 2 * dhc * (pf * rf) * P.mass;
            }
        });
        drho[j] = s;
        pres[j] = eos_p(P, rho[j]);
    }
}

// ExtForce.h:15-61 (repulsive LJ walls)
template <int D>
inline V<D> ext_part(const oracle_params& P, double dist, V<D> rj, V<D> ri, int32_t tj, int32_t ti) {
    if (P.boundary_treatment == 1 && tj > 0 && ti < 0) {
        const double r0 = 2 * P.hsml;
        const double Dc = 5 * P.g * P.depth;
        if (dist < r0) {
            double rn = r0 / dist;
            double f = Dc * (mulpow(rn, 12) - mulpow(rn, 4)) / sq(dist);
            return (rj - ri) * f;
        }
    }
    return zero<D>();
}

// InternalForce.h:24-48
inline double art_pressure(const oracle_params& P, const Kern& K, double wij, double pi, double pj,
                           double rhoi, double rhoj) {
    const double wdelta = K.w(P.delta, P.artificial_pressure_skf);
    const bool ni = pi < 0, nj = pj < 0;
    if (!(ni || nj)) return 0;
    double fij = wij / wdelta;
    double api = ni ? (-P.artificial_pressure_coef * pi / sq(rhoi)) : 0;
    double apj = nj ? (-P.artificial_pressure_coef * pj / sq(rhoj)) : 0;
    return (api + apj) * std::pow(fij, P.artificial_pressure_index);
}

// InternalForce.h:53-141
template <int D>
inline V<D> intf_part(const oracle_params& P, const Kern& K, V<D> diff, double dist, double pj, double pi,
                      double rhoj, double rhoi) {
    V<D> g = K.dwdr<D>(dist, diff, P.intf_skf);
    double pf;
    if (P.intf_sph_approximation == 1) {
        pf = (pj + pi) / (rhoj * rhoi);
    } else {
        pf = pi / sq(rhoi) + pj / sq(rhoj);
    }
    if (P.artificial_pressure)
        pf += art_pressure(P, K, K.w(dist, P.artificial_pressure_skf), pi, pj, rhoi, rhoj);
    V<D> h = -g * pf;
    return -h * P.mass;
}

// ArtificialViscosity.h:29-66
template <int D>
inline V<D> av_part(const oracle_params& P, const Kern& K, V<D> rj, V<D> ri, V<D> vj, V<D> vi, double rhoj,
                    double rhoi, double* arvmu) {
    V<D> dv = vj - vi;
    V<D> dr = rj - ri;
    double vr = dot(dv, dr);
    double rr = len2(dr);
    if (vr < 0) {
        const double etq = (double)0.1f;
        double mu = P.hsml * vr / (rr + sq(P.hsml * etq));
        if (P.dt_correction_method == 2) *arvmu = std::max(*arvmu, std::fabs(mu));
        double rhoij = 0.5 * (rhoi + rhoj);
        double piv = (P.artificial_bulk_visc * mu - P.artificial_shear_visc * P.eos_sound_vel) * mu / rhoij;
        V<D> diff = ri - rj;
        V<D> g = K.dwdr<D>(len(diff), diff, P.artificial_viscosity_skf);
        return g * piv * P.mass;
    }
    return zero<D>();
}

// AverageVelocity.h:20-37
template <int D>
inline V<D> xsph_part(const oracle_params& P, const Kern& K, double dist, int32_t ti, V<D> vj, V<D> vi,
                      double rhoj, double rhoi) {
    if (ti > 0) {
        double w = K.w(dist, P.average_velocity_skf);
        V<D> dv = vi - vj;
        return dv * P.mass / (rhoi + rhoj) * w * 2 * P.average_velocity_coef;
    }
    return zero<D>();
}

// UpdateAcceleration.cpp:33-116
template <int D>
void update_acceleration_t(const oracle_params& P, const int32_t* itype, const double* r, const double* v,
                           const uint32_t* nbr, const double* rho, const double* pres, double* a, double* av,
                           double* arvmu_out, double* amax, double* mumax) {
    Kern K{P};
    double am = 0, mm = 0;
#pragma omp parallel for schedule(static) reduction(max : am, mm)
    for (int64_t jj = 0; jj < (int64_t)P.ntotal; ++jj) {
        const uint32_t j = (uint32_t)jj;
        double arvmu = 0;
        V<D> avt = zero<D>();
        V<D> at = zero<D>();
        at.y = -P.g;
        V<D> rj = V<D>::load(r, j), vj = V<D>::load(v, j);
        for_pairs<D>(P, nbr, j, [&](uint32_t i) {
            V<D> ri = V<D>::load(r, i), vi = V<D>::load(v, i);
            V<D> diff = ri - rj;
            double dist = len(diff);
            at += ext_part<D>(P, dist, rj, ri, itype[j], itype[i]);
            at += intf_part<D>(P, K, diff, dist, pres[j], pres[i], rho[j], rho[i]);
            if (P.artificial_viscosity) at += av_part<D>(P, K, rj, ri, vj, vi, rho[j], rho[i], &arvmu);
            if (P.average_velocity) avt += xsph_part<D>(P, K, dist, itype[i], vj, vi, rho[j], rho[i]);
        });
        at.store(a, j);
        if (av) avt.store(av, j);
        if (arvmu_out) arvmu_out[j] = arvmu;
        am = std::max(am, len(at));
        mm = std::max(mm, arvmu);
    }
    if (amax) *amax = am;
    if (mumax) *mumax = mm;
}

// ---------------------------------------------------------------- wave maker
// WaveMaker.cpp:34-82 (first order), cl/TimeIntegration.cl:196-245 (second
// order coefficient as the product e2_1 * e2_2, matching the documentation),
// plus the Fourier-series piston x(t) = sum a_k sin(k w t) + b_k cos(k w t).
// A segmented Fourier piston (3D) moves paddle `seg` with its own harmonics
// (paddle-major coefficients).
double piston_velocity(const oracle_params& P, double time, int seg = 0) {
    if (P.nwm == 6) {
        const double* a = P.piston_a + seg * P.piston_nharm;
        const double* b = P.piston_b + seg * P.piston_nharm;
        double v = 0;
        for (int k = 1; k <= P.piston_nharm; ++k) {
            double kw = k * P.piston_omega;
            v += kw * (a[k - 1] * std::cos(kw * time) - b[k - 1] * std::sin(kw * time));
        }
        return v;
    }
    const double delta = P.nwm_phase - P.nwm_freq * P.nwm_time_start;
    const double omega = P.nwm_freq;
    const double H = P.nwm_wave_magnitude;
    const double kd = P.nwm_wave_number * P.depth;
    const double m1 = 2 * sq(std::sinh(kd)) / (std::sinh(kd) * std::cosh(kd) + kd);
    const double S0 = H / m1;
    double vx = 0.5 * S0 * omega * std::cos(omega * time + delta);
    if (P.nwm == 3) {
        double e2 = (sq(H) / (32 * P.depth)) * (3 * std::cosh(kd) / cb(std::sinh(kd)) - 2 / m1);
        vx += 2 * omega * e2 * std::cos(2 * omega * time + 2 * delta);
    }
    return vx;
}

// WaveMaker.cpp:84-134 (solitary wave, Rayleigh approximation)
double solitary_velocity(const oracle_params& P, double time) {
    const double H = P.nwm_wave_magnitude, h = P.depth, g = P.g;
    const double k = std::sqrt(3 * H / (4 * sq(h) * (H + h)));
    const double c = std::sqrt(g * (H + h));
    const double Tf = 2 / (k * c) * (3.8 + H / h);
    const double tau = k * c * (time - Tf);
    const double u = std::tanh(tau);
    const double top = k * H * (h + H - H * sq(u)) + 2 * k * sq(H * u);
    const double bottom = sq(k * (h + H - H * sq(u)));
    return top / bottom * (1 - sq(std::tanh(tau))) * (k * c);
}

}  // namespace

// ==================================================================== context
struct oracle_ctx {
    oracle_params P;
    std::vector<double> r, v, rho, p, drho, a, av, rho_pred, v_pred, arvmu, rho_old;
    std::vector<int32_t> itype;
    std::vector<uint32_t> nbr, grid, start;
    std::vector<uint64_t> key;
    double time = 0, dt = 0;
    uint64_t step = 0;
    uint32_t outside = 0;
};

template <int D>
static int32_t run_steps(oracle_ctx* c, uint32_t nsteps) {
    oracle_params& P = c->P;
    const uint32_t n = P.ntotal;
    const bool cont = P.density_treatment != 0;
    const double c0 = P.eos_c;
    const double never_ending = (P.hsml / c0) * 1.E-6;
    for (uint32_t s = 0; s < nsteps; ++s) {
        // predict_half_step, PredictHalfStep.cpp:4-23
        for (uint32_t i = 0; i < n; ++i) {
            if (cont) c->rho_pred[i] = c->rho[i] + c->drho[i] * c->dt * 0.5;
            for (int k = 0; k < D; ++k)
                c->v_pred[D * i + k] = c->v[D * i + k] + c->a[D * i + k] * c->dt * 0.5;
        }
        // grid_find, GridFind.cpp:215-236
        int32_t st = grid_find_t<D>(P, c->r.data(), c->itype.data(), c->nbr.data(), c->grid, c->start, c->key);
        if (st) return st;
        // density, TimeIntegration.cpp:74-83
        // the renormalised summation on every density_normalization_every-th step
        const uint32_t every = P.density_normalization_every ? P.density_normalization_every : 1;
        const bool renorm = !cont && P.density_normalization && (c->step + 1) % every == 0;
        if (cont) {
            density_con_t<D>(P, c->r.data(), c->v.data(), c->nbr.data(), c->rho.data(), c->drho.data(), c->p.data());
        } else if (renorm) {
            c->rho_old = c->rho;
            density_sum_t<D>(P, c->r.data(), c->nbr.data(), c->rho.data(), c->p.data(), c->rho_old.data());
        } else {
            density_sum_t<D>(P, c->r.data(), c->nbr.data(), c->rho.data(), c->p.data());
        }
        // update_acceleration + update_dt, UpdateAcceleration.cpp:15-124
        double amax = 0, mumax = 0;
        update_acceleration_t<D>(P, c->itype.data(), c->r.data(), c->v_pred.data(), c->nbr.data(),
                                 cont ? c->rho_pred.data() : c->rho.data(), c->p.data(), c->a.data(),
                                 c->av.data(), c->arvmu.data(), &amax, &mumax);
        if (P.dt_correction_method == 2) {
            if (P.dt_dynamic_mode == 0) { amax = 0; mumax = 0; }
            double dta = std::sqrt(P.hsml / amax);
            double dtm = P.hsml / (c0 + mumax);
            c->dt = P.CFL_coef * std::min(dta, dtm);
            if (c->dt <= never_ending) return 2;
        }
        // wave maker (boundary motion only): CLTimeIntegration.cpp:369-413
        if (P.boundary_motion && P.nwm && c->time >= P.nwm_time_start &&
            !(P.nwm_particles_start < P.nfluid || P.nwm_particles_end < P.nfluid ||
              P.nwm_particles_end < P.nwm_particles_start || P.nwm_particles_end > n)) {
            if (P.nwm == 4) {
                for (uint32_t i = P.nwm_particles_start; i < P.nwm_particles_end; ++i) c->itype[i] = 0;
                P.nwm = 0;
            } else {
                // cl/TimeIntegration.cl:231-244: v = vx * dir; an undefined
                // direction gives v = 0; the solitary piston always moves along x
                double vx = P.nwm == 5 ? solitary_velocity(P, c->time) : piston_velocity(P, c->time);
                int axis = P.nwm == 5 ? 0 : P.nwm_direction - 1;
                const int segs = P.piston_segments > 1 && D == 3 && P.nwm == 6 ? P.piston_segments : 1;
                const double seg_w = segs > 1 ? segs / (P.piston_seg_zmax - P.piston_seg_zmin) : 0.0;
                for (uint32_t i = P.nwm_particles_start; i < P.nwm_particles_end; ++i) {
                    if (segs > 1) {  // paddle of the particle: floor((z - zmin) * segs / (zmax - zmin)), clamped
                        const double u = std::floor((c->r[D * i + D - 1] - P.piston_seg_zmin) * seg_w);
                        const int g = u < 0.0 ? 0 : (u >= segs ? segs - 1 : (int)u);
                        vx = piston_velocity(P, c->time, g);
                    }
                    for (int k = 0; k < D; ++k) c->v[D * i + k] = 0;
                    if (axis >= 0 && axis < D) {
                        c->v[D * i + axis] = vx;
                    } else if (axis == 3 && D == 3) {  // radial in the x-z plane
                        double x = c->r[D * i], z = c->r[D * i + 2];
                        double nrm = std::sqrt(x * x + z * z);
                        c->v[D * i] = vx * (x / nrm);
                        c->v[D * i + 2] = vx * (z / nrm);
                    }
                }
            }
        }
        // whole_step, WholeStep.cpp:19-61 (+ cl/TimeIntegration.cl:190-193)
        const double rdt = c->dt;
        const double vdt = c->step ? c->dt : c->dt * 0.5;
        for (uint32_t i = 0; i < n; ++i) {
            if (cont) c->rho[i] = c->rho[i] + c->drho[i] * vdt;
            if (c->itype[i] > 0) {
                for (int k = 0; k < D; ++k) c->v[D * i + k] += c->a[D * i + k] * vdt;
                if (P.average_velocity)
                    for (int k = 0; k < D; ++k) c->v[D * i + k] += c->av[D * i + k];
                double nr[3];
                for (int k = 0; k < D; ++k) nr[k] = c->r[D * i + k] + c->v[D * i + k] * rdt;
                if (P.consistency_treatment == 2) {
                    bool inside = nr[0] < P.x_maxgeom && nr[0] > P.x_mingeom && nr[1] < P.y_maxgeom &&
                                  nr[1] > P.y_mingeom && (D == 2 || (nr[2] < P.z_maxgeom && nr[2] > P.z_mingeom));
                    if (inside)
                        for (int k = 0; k < D; ++k) c->r[D * i + k] = nr[k];
                    else
                        c->itype[i] = 0;
                } else {
                    for (int k = 0; k < D; ++k) c->r[D * i + k] = nr[k];
                }
            } else if (P.boundary_motion && c->itype[i] < 0) {
                for (int k = 0; k < D; ++k) c->r[D * i + k] += c->v[D * i + k] * rdt;
            }
        }
        c->time += c->dt;
        c->step++;
        // per-step in-bounds check, ConsistencyCheck.cpp:162-189
        uint32_t out = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const double* q = &c->r[D * i];
            bool ok = q[0] < P.x_maxgeom && q[0] > P.x_mingeom && std::isfinite(q[0]) && q[1] < P.y_maxgeom &&
                      q[1] > P.y_mingeom && std::isfinite(q[1]);
            if (D == 3) ok = ok && q[2] < P.z_maxgeom && q[2] > P.z_mingeom && std::isfinite(q[2]);
            if (!ok) ++out;
        }
        c->outside = out;
        if (P.consistency_check && out > 0 && P.consistency_treatment == 1) return 3;
    }
    return 0;
}

template <int D, class F>
static void per_pair(const oracle_params* p, const uint32_t* nb, F&& f) {
    for (uint32_t j = 0; j < p->ntotal; ++j) for_pairs<D>(*p, nb, j, [&](uint32_t i) { f(j, i); });
}

// FuncAtPoint (tools/FuncAtPoint/lib/src/fatp.cpp:121-144 calculate, :14-58
// findValue): the layer's grid_find and density_sum, then at each point the
// sum over every particle of the 3^dim cells around the point's cell (the
// grid holds all particles, dead ones too) with |r_i - x|^2 < cell^2 of
// f_i W_cubic(|r_i - x|) m / rho_i, in the cells' order.
template <int D>
static int32_t probe_t(const oracle_params& P, const double* r, const int32_t* itype, const double* f, uint32_t stride,
                uint32_t comp, const double* pts, uint32_t np, double* out) {
    std::vector<uint32_t> grid, start, nbr((size_t)P.max_neighbours * P.ntotal);
    std::vector<uint64_t> key;
    const int32_t st = grid_find_t<D>(P, r, itype, nbr.data(), grid, start, key);
    if (st != 0) return st;
    std::vector<double> rho(P.ntotal), pres(P.ntotal);
    density_sum_t<D>(P, r, nbr.data(), rho.data(), pres.data());
    Kern K{P};
    const double max_d2 = sq(cell_size(P));
    const uint64_t tsz = table_size(P);
    const int ncell = D == 3 ? 27 : 9;
    for (uint32_t g = 0; g < np; ++g) {
        const V<D> x = V<D>::load(pts, g);
        uint64_t cells[27];
        neighbour_cells(D, cell_of<D>(P, x), cells);
        double val = 0;
        for (int k = 0; k < ncell; ++k) {
            const uint64_t c = cells[k];
            if (c == kInvalidCell || c + 1 >= tsz) continue;
            for (uint32_t gi = start[c]; gi < start[c + 1]; ++gi) {
                const uint32_t i = grid[gi];
                const double d2 = len2(V<D>::load(r, i) - x);
                if (d2 < max_d2) val += f[(size_t)i * stride + comp] * K.w(std::sqrt(d2), 1) * P.mass / rho[i];
            }
        }
        out[g] = val;
    }
    return 0;
}

extern "C" {

double oracle_kernel_w(const oracle_params* p, double dist, int32_t skf) { return Kern{*p}.w(dist, skf); }

void oracle_kernel_dwdr(const oracle_params* p, double dist, const double* diff, int32_t skf, double* out) {
    Kern K{*p};
    if (p->dim == 3) K.dwdr<3>(dist, V<3>(diff[0], diff[1], diff[2]), skf).store(out, 0);
    else K.dwdr<2>(dist, V<2>(diff[0], diff[1]), skf).store(out, 0);
}

double oracle_eos_p(const oracle_params* p, double rho) { return eos_p(*p, rho); }

// EOS.cpp:36-44
double oracle_eos_rho(const oracle_params* p, double pressure) {
    double val = pressure / eos_B(*p) + 1;
    if (val >= 0) return p->rho0 * std::pow(val, 1.0 / 7);
    return NAN;
}

uint32_t oracle_morton2(uint32_t x, uint32_t y) { return split2(x) | (split2(y) << 1); }
void oracle_morton2_decode(uint32_t key, uint32_t* x, uint32_t* y) { *x = join2(key); *y = join2(key >> 1); }
uint32_t oracle_morton3(uint32_t x, uint32_t y, uint32_t z) { return split3(x) | (split3(y) << 1) | (split3(z) << 2); }
void oracle_morton3_decode(uint32_t key, uint32_t* x, uint32_t* y, uint32_t* z) {
    *x = join3(key); *y = join3(key >> 1); *z = join3(key >> 2);
}
uint64_t oracle_cell_coord(const oracle_params* p, double r, double rmin) { return coord_of(*p, r, rmin); }
uint64_t oracle_cell_idx(const oracle_params* p, const double* r) {
    return p->dim == 3 ? cell_of<3>(*p, V<3>::load(r, 0)) : cell_of<2>(*p, V<2>::load(r, 0));
}
void oracle_neighbouring_cells(const oracle_params* p, uint64_t idx, uint64_t* cells) {
    neighbour_cells(p->dim, idx, cells);
}
uint64_t oracle_count_cells(const oracle_params* p, int32_t* status) { return count_cells(*p, status); }

static void set_threads(const oracle_params* p) {
#ifdef _OPENMP
    if (p->num_threads > 0) omp_set_num_threads(p->num_threads);
#endif
}

int32_t oracle_grid_find(const oracle_params* p, const double* r, const int32_t* itype, uint32_t* neighbours) {
    set_threads(p);
    std::vector<uint32_t> grid, start;
    std::vector<uint64_t> key;
    return p->dim == 3 ? grid_find_t<3>(*p, r, itype, neighbours, grid, start, key)
                       : grid_find_t<2>(*p, r, itype, neighbours, grid, start, key);
}

void oracle_density_sum(const oracle_params* p, const double* r, const uint32_t* nb, double* rho, double* pres) {
    set_threads(p);
    if (p->dim == 3) density_sum_t<3>(*p, r, nb, rho, pres);
    else density_sum_t<2>(*p, r, nb, rho, pres);
}

void oracle_density_sum_normalized(const oracle_params* p, const double* r, const uint32_t* nb, const double* rho_old,
                                   double* rho, double* pres) {
    set_threads(p);
    if (p->dim == 3) density_sum_t<3>(*p, r, nb, rho, pres, rho_old);
    else density_sum_t<2>(*p, r, nb, rho, pres, rho_old);
}

void oracle_density_con(const oracle_params* p, const double* r, const double* v, const uint32_t* nb,
                        const double* rho, double* drho, double* pres) {
    set_threads(p);
    if (p->dim == 3) density_con_t<3>(*p, r, v, nb, rho, drho, pres);
    else density_con_t<2>(*p, r, v, nb, rho, drho, pres);
}

int32_t oracle_probe(const oracle_params* p, const double* r, const int32_t* itype, const double* field,
                     uint32_t stride, uint32_t comp, const double* points, uint32_t npoints, double* out) {
    set_threads(p);
    return p->dim == 3 ? probe_t<3>(*p, r, itype, field, stride, comp, points, npoints, out)
                       : probe_t<2>(*p, r, itype, field, stride, comp, points, npoints, out);
}

// testExternalForceOMP.cpp:33-55: start at (0, -g) and add each pair term
void oracle_term_external(const oracle_params* p, const double* r, const int32_t* itype, const uint32_t* nb,
                          double* out) {
    const int D = p->dim;
    for (uint32_t j = 0; j < p->ntotal; ++j) {
        for (int k = 0; k < D; ++k) out[D * j + k] = 0;
        out[D * j + 1] = -p->g;
    }
    if (D == 3) {
        per_pair<3>(p, nb, [&](uint32_t j, uint32_t i) {
            V<3> rj = V<3>::load(r, j), ri = V<3>::load(r, i);
            (V<3>::load(out, j) + ext_part<3>(*p, len(rj - ri), rj, ri, itype[j], itype[i])).store(out, j);
        });
    } else {
        per_pair<2>(p, nb, [&](uint32_t j, uint32_t i) {
            V<2> rj = V<2>::load(r, j), ri = V<2>::load(r, i);
            (V<2>::load(out, j) + ext_part<2>(*p, len(rj - ri), rj, ri, itype[j], itype[i])).store(out, j);
        });
    }
}

// testInternalForceOMP.cpp:287-323
void oracle_term_internal(const oracle_params* p, const double* r, const uint32_t* nb, const double* pres,
                          const double* rho, double* out) {
    Kern K{*p};
    std::memset(out, 0, sizeof(double) * p->dim * p->ntotal);
    if (p->dim == 3) {
        per_pair<3>(p, nb, [&](uint32_t j, uint32_t i) {
            V<3> d = V<3>::load(r, i) - V<3>::load(r, j);
            (V<3>::load(out, j) + intf_part<3>(*p, K, d, len(d), pres[j], pres[i], rho[j], rho[i])).store(out, j);
        });
    } else {
        per_pair<2>(p, nb, [&](uint32_t j, uint32_t i) {
            V<2> d = V<2>::load(r, i) - V<2>::load(r, j);
            (V<2>::load(out, j) + intf_part<2>(*p, K, d, len(d), pres[j], pres[i], rho[j], rho[i])).store(out, j);
        });
    }
}

// testArtificialViscosityOMP.cpp:149-171
void oracle_term_art_visc(const oracle_params* p, const double* r, const double* v, const uint32_t* nb,
                          const double* rho, double* out, double* mu) {
    Kern K{*p};
    std::memset(out, 0, sizeof(double) * p->dim * p->ntotal);
    std::memset(mu, 0, sizeof(double) * p->ntotal);
    if (p->dim == 3) {
        per_pair<3>(p, nb, [&](uint32_t j, uint32_t i) {
            (V<3>::load(out, j) + av_part<3>(*p, K, V<3>::load(r, j), V<3>::load(r, i), V<3>::load(v, j),
                                             V<3>::load(v, i), rho[j], rho[i], &mu[j])).store(out, j);
        });
    } else {
        per_pair<2>(p, nb, [&](uint32_t j, uint32_t i) {
            (V<2>::load(out, j) + av_part<2>(*p, K, V<2>::load(r, j), V<2>::load(r, i), V<2>::load(v, j),
                                             V<2>::load(v, i), rho[j], rho[i], &mu[j])).store(out, j);
        });
    }
}

// testAverageVelocityOMP.cpp:282-305 (only j < jmax contribute)
void oracle_term_average_velocity(const oracle_params* p, const double* r, const double* v, const int32_t* itype,
                                  const uint32_t* nb, const double* rho, uint32_t jmax, double* out) {
    Kern K{*p};
    std::memset(out, 0, sizeof(double) * p->dim * p->ntotal);
    if (p->dim == 3) {
        per_pair<3>(p, nb, [&](uint32_t j, uint32_t i) {
            if (j >= jmax) return;
            double d = len(V<3>::load(r, j) - V<3>::load(r, i));
            (V<3>::load(out, j) + xsph_part<3>(*p, K, d, itype[i], V<3>::load(v, j), V<3>::load(v, i), rho[j],
                                               rho[i])).store(out, j);
        });
    } else {
        per_pair<2>(p, nb, [&](uint32_t j, uint32_t i) {
            if (j >= jmax) return;
            double d = len(V<2>::load(r, j) - V<2>::load(r, i));
            (V<2>::load(out, j) + xsph_part<2>(*p, K, d, itype[i], V<2>::load(v, j), V<2>::load(v, i), rho[j],
                                               rho[i])).store(out, j);
        });
    }
}

void oracle_update_acceleration(const oracle_params* p, const int32_t* itype, const double* r, const double* v,
                                const uint32_t* nb, const double* rho, const double* pres, double* a, double* av,
                                double* arvmu, double* amax, double* mumax) {
    set_threads(p);
    if (p->dim == 3) update_acceleration_t<3>(*p, itype, r, v, nb, rho, pres, a, av, arvmu, amax, mumax);
    else update_acceleration_t<2>(*p, itype, r, v, nb, rho, pres, a, av, arvmu, amax, mumax);
}

oracle_ctx* oracle_create(const oracle_params* p) {
    if (!p || (p->dim != 2 && p->dim != 3) || p->max_neighbours < 2) return nullptr;
    oracle_ctx* c = new oracle_ctx();
    c->P = *p;
    const size_t n = p->ntotal, D = p->dim;
    c->r.assign(D * n, 0); c->v.assign(D * n, 0); c->a.assign(D * n, 0); c->av.assign(D * n, 0);
    c->v_pred.assign(D * n, 0);
    c->rho.assign(n, 0); c->p.assign(n, 0); c->drho.assign(n, 0); c->rho_pred.assign(n, 0); c->arvmu.assign(n, 0);
    c->itype.assign(n, 0);
    c->nbr.assign((size_t)p->max_neighbours * n, 0);
    c->dt = p->dt;
    return c;
}

void oracle_destroy(oracle_ctx* c) { delete c; }

int32_t oracle_set_state(oracle_ctx* c, const double* r, const double* v, const double* rho, const double* pres,
                         const int32_t* itype, double time, double dt) {
    const size_t n = c->P.ntotal, D = c->P.dim;
    std::copy(r, r + D * n, c->r.begin());
    std::copy(v, v + D * n, c->v.begin());
    std::copy(rho, rho + n, c->rho.begin());
    if (pres) std::copy(pres, pres + n, c->p.begin());
    std::copy(itype, itype + n, c->itype.begin());
    std::fill(c->a.begin(), c->a.end(), 0.0);
    std::fill(c->drho.begin(), c->drho.end(), 0.0);
    c->time = time;
    c->dt = dt;
    c->step = 0;
    return 0;
}

int32_t oracle_step(oracle_ctx* c, uint32_t nsteps) {
    set_threads(&c->P);
    return c->P.dim == 3 ? run_steps<3>(c, nsteps) : run_steps<2>(c, nsteps);
}

void oracle_get_state(const oracle_ctx* c, double* r, double* v, double* rho, double* pres, int32_t* itype,
                      double* time, double* dt, uint64_t* step) {
    if (r) std::copy(c->r.begin(), c->r.end(), r);
    if (v) std::copy(c->v.begin(), c->v.end(), v);
    if (rho) std::copy(c->rho.begin(), c->rho.end(), rho);
    if (pres) std::copy(c->p.begin(), c->p.end(), pres);
    if (itype) std::copy(c->itype.begin(), c->itype.end(), itype);
    if (time) *time = c->time;
    if (dt) *dt = c->dt;
    if (step) *step = c->step;
}

void oracle_get_fields(const oracle_ctx* c, double* a, double* drho, double* av, double* arvmu) {
    if (a) std::copy(c->a.begin(), c->a.end(), a);
    if (drho) std::copy(c->drho.begin(), c->drho.end(), drho);
    if (av) std::copy(c->av.begin(), c->av.end(), av);
    if (arvmu) std::copy(c->arvmu.begin(), c->arvmu.end(), arvmu);
}

const uint32_t* oracle_neighbours_ptr(const oracle_ctx* c) { return c->nbr.data(); }
uint32_t oracle_last_outside(const oracle_ctx* c) { return c->outside; }
double oracle_piston_velocity(const oracle_params* p, double time) {
    return p->nwm == 5 ? solitary_velocity(*p, time) : piston_velocity(*p, time);
}

}  // extern "C"
