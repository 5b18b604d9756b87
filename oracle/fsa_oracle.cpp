// fsa_oracle.cpp -- CPU restatement of the FluidSim per-step particle
// pipeline.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP path
// is compared against, and the CPU baseline timed beside it.  Nothing in the
// product (fluidsim_amd/) links, loads or calls it.
//
// Templated on the arithmetic type T (float = the reference's FP32, double =
// the FP64 precision oracle) and on the neighbour traversal:
//   CANONICAL  fluid and wall particles sorted by (cell id, eighth-cell z
//              index inside the cell, .dat index) -- the eighth-cell index
//              is (int)(8 t) - 8 (int)t with t = (z - cell_min_z) /
//              interval_z -- and solid particles by (cell id, .dat index), the key order of
//              the GPU's cell build; for a particle in cell
//              (cx,cy,cz) the partners are walked kind by kind (fluid, solid,
//              virt), and for each kind the 27 neighbour cells as 9 columns
//              (ox,oy ascending), each column as the contiguous z-run
//              cz-1..cz+1.  This is the order the HIP kernels use.
//   EXACT      the reference's per-cell singly linked lists: host build of
//              src/sim.cu:256-302 (append, ascending index), per-step rebuild
//              of src/pair.cu:72-207 (walk previous lists of the 27
//              neighbours in the fixed order of src/sim.cu:166-230, append
//              matches; a particle that moved more than one cell is lost),
//              and traversal per neighbour cell n=0..26 of fluid, solid,
//              virt lists as in src/fluid.cu:188-227.
// The two visit the same neighbour sets; only the summation order differs.
//
// Every stage follows the reference file:line named at the function.
// Candidate distances are computed as fma(dz,dz,fma(dy,dy,dx*dx)) so that the
// candidate test is bit-reproducible between this code and the GPU.
// Build with -ffp-contract=off so that no other contraction happens.

#include "fluidsim_amd.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { KF = 0, KS = 1, KV = 2 };
enum { CANONICAL = 0, EXACT = 1 };

// constants of include/global.h:4-10
constexpr float kEpsForSum = 0.1f;
constexpr float kEpsFloat = 1.0e-6f;
constexpr float kLargeFloat = 1.0e6f;

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> &operator+=(V3<T> &a, V3<T> b) { a = a + b; return a; }
template <class T> inline V3<T> &operator-=(V3<T> &a, V3<T> b) { a = a - b; return a; }
// glm::dot: (x*x' + y*y') + z*z'
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// squared length in the fused form fma(z,z,fma(y,y,x*x)); every glm::length2
// (and the dot(v,v) inside glm::normalize / glm::length) of the reference is
// evaluated this way here and on the GPU, so candidate tests agree bit for bit
template <class T> inline T dist2(V3<T> d) { return std::fma(d.z, d.z, std::fma(d.y, d.y, d.x * d.x)); }
template <class T> inline T length2(V3<T> a) { return dist2(a); }
// glm::normalize = v * inversesqrt(dot(v, v))
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / std::sqrt(length2(a))); }

// 3x3 matrix, m[r][c] = row r, column c (glm stores the transpose of this)
template <class T> struct M3 {
    T m[3][3];
};
template <class T> inline M3<T> mzero() { M3<T> r; for (auto &row : r.m) for (auto &v : row) v = T(0); return r; }
template <class T> inline M3<T> mident() { M3<T> r = mzero<T>(); r.m[0][0] = r.m[1][1] = r.m[2][2] = T(1); return r; }
template <class T> inline M3<T> operator+(const M3<T> &a, const M3<T> &b) {
    M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j]; return r;
}
template <class T> inline M3<T> operator-(const M3<T> &a, const M3<T> &b) {
    M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j]; return r;
}
template <class T> inline M3<T> operator*(const M3<T> &a, T s) {
    M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s; return r;
}
template <class T> inline M3<T> operator*(T s, const M3<T> &a) {
    M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j]; return r;
}
template <class T> inline M3<T> operator*(const M3<T> &a, const M3<T> &b) {
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
template <class T> inline V3<T> operator*(const M3<T> &a, V3<T> v) {
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
            a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
template <class T> inline M3<T> transpose(const M3<T> &a) {
    M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i]; return r;
}
// glm::outerProduct(c, r): element (row i, col j) = c_i * r_j
template <class T> inline M3<T> outer(V3<T> c, V3<T> r) {
    const T cv[3] = {c.x, c.y, c.z}, rv[3] = {r.x, r.y, r.z};
    M3<T> o; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) o.m[i][j] = cv[i] * rv[j]; return o;
}
// determinant by cofactor expansion along the first row
template <class T> inline T det(const M3<T> &a) {
    const auto &m = a.m;
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
// inverse = adjugate / determinant
template <class T> inline M3<T> inverse(const M3<T> &a) {
    const auto &m = a.m;
    T c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    T c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    T c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    T inv_det = T(1) / (m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02);
    M3<T> r;
    r.m[0][0] = c00 * inv_det;
    r.m[1][0] = c01 * inv_det;
    r.m[2][0] = c02 * inv_det;
    r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv_det;
    r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv_det;
    r.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv_det;
    r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv_det;
    r.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv_det;
    r.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv_det;
    return r;
}

// parameters of fsa_params converted to T
template <class T> struct Par {
    int nf, ns, nv;
    V3<T> g;
    T cs, rho0, rho0_inv, gamma, visc, rho2p, p_min, rho_min, rho_max;
    bool viscid;
    T rho_s, lambda, mu, cs_s, alpha, beta;
    T dmin[3], dmax[3], cmin[3], inv[3];
    int cn[3], ncell;
    T dt;
    int sub;
    T h, h2, h_inv2, R2, c1, c2, kd;
    bool compensated;  // FSA_FORM_COMPENSATED: Kahan updates of x and u
};

// v += d with compensated (Kahan) summation, c carrying the rounding error
// (Kahan, "Further remarks on reducing truncation errors", CACM 8(1), 1965)
template <class T> inline T kahan_add(T v, T d, T &c) {
    const T y = d - c;
    const T t = v + y;
    c = (t - v) - y;
    return t;
}

template <class T> Par<T> convert(const fsa_params &p) {
    Par<T> q;
    q.nf = p.n_fluid;
    q.ns = p.n_solid;
    q.nv = p.n_virt;
    q.g = {T(p.gravity[0]), T(p.gravity[1]), T(p.gravity[2])};
    q.cs = p.fluid_sound_speed;
    q.rho0 = p.fluid_reference_density;
    q.rho0_inv = p.fluid_reference_density_inverse;
    q.gamma = p.fluid_gamma;
    q.visc = p.fluid_viscosity;
    q.rho2p = p.fluid_coefficient_rho2p;
    q.p_min = p.fluid_min_pressure;
    q.rho_min = p.fluid_min_density;
    q.rho_max = p.fluid_max_density;
    q.viscid = p.fluid_is_viscid != 0;
    q.rho_s = p.solid_reference_density;
    q.lambda = p.solid_lambda;
    q.mu = p.solid_nu;
    q.cs_s = p.solid_sound_speed;
    q.alpha = p.solid_artificial_viscosity[0];
    q.beta = p.solid_artificial_viscosity[1];
    for (int a = 0; a < 3; ++a) {
        q.dmin[a] = p.domain_min[a];
        q.dmax[a] = p.domain_max[a];
        q.cmin[a] = p.cell_min[a];
        q.inv[a] = p.interval_inv[a];
        q.cn[a] = p.cell_number[a];
    }
    q.ncell = p.cell_number_total;
    q.dt = p.dt;
    q.sub = p.solid_sub_step;
    q.h = p.smoothing_length;
    q.h2 = p.smoothing_length_square;
    q.h_inv2 = p.smoothing_length_inverse_square;
    q.R2 = p.impact_distance_square;
    q.compensated = (p.precision_form & FSA_FORM_COMPENSATED) != 0;
    q.c1 = p.kernel_coefficient_1;
    q.c2 = p.kernel_coefficient_2;
    q.kd = p.kernel_differential;
    return q;
}

struct OracleBase {
    virtual ~OracleBase() {}
    std::string err;
    virtual int upload_fluid(const float *, const float *, const float *, const float *) = 0;
    virtual int upload_solid(const float *, const float *, const float *, const int32_t *) = 0;
    virtual int upload_virt(const float *, const int32_t *) = 0;
    virtual int download_fluid(double *, double *, double *, double *, double *, double *) = 0;
    virtual int download_solid(double *, double *, double *, double *, double *, double *, double *, double *,
                               double *) = 0;
    virtual int download_virt(double *, double *, double *, double *) = 0;
    virtual void particle_cell_update() = 0;
    virtual void fill_cell() = 0;
    virtual void couple_calculation() = 0;
    virtual void virt_update() = 0;
    virtual void fluid_calculation() = 0;
    virtual void solid_calculation() = 0;
    virtual void solid_deformation() = 0;
    virtual void solid_force() = 0;
    virtual void solid_update() = 0;
    virtual int cell_ids(int kind, int32_t *out) = 0;
    virtual int cell_start(int kind, int32_t *out) = 0;
    virtual int sorted_ids(int kind, int32_t *out) = 0;
    virtual int neighbors(int kind_i, int i, unsigned mask, int reference_metric, int32_t *out_kind, int32_t *out_j,
                          int cap) = 0;
    virtual int brute_neighbors(int kind_i, int i, unsigned mask, int reference_metric, int32_t *out_kind,
                                int32_t *out_j, int cap) = 0;
    virtual int64_t lost_count() = 0;
    virtual void pair_counts(int brute, int32_t *out) = 0;
    void step() {
        particle_cell_update();
        fill_cell();
        couple_calculation();
        virt_update();
        fluid_calculation();
        solid_calculation();
    }
};

template <class T> struct Oracle : OracleBase {
    Par<T> P;
    int mode;
    // fluid (Fluid_List, include/particle.cuh:20-32)
    std::vector<V3<T>> fx, fu, fa;
    std::vector<T> frho, fp, fdrho;
    std::vector<V3<T>> fcx, fcu, scx, scu;  // carried rounding errors of x and u (compensated form)
    // solid (Solid_List, include/particle.cuh:91-102)
    std::vector<V3<T>> sx, su, sX, sn, sa, sac;
    std::vector<int> stype;
    std::vector<T> sp;
    std::vector<M3<T>> sF, sP, sL;
    // virt (Virt_List, include/particle.cuh:171-177)
    std::vector<V3<T>> vx, vn, vu;
    std::vector<T> vrho, vp;
    std::vector<int> vtype;
    // cell id and eighth-cell z index per particle and kind
    std::vector<int> cell[3], sub[3];
    static constexpr int kSubZ = 8;
    // canonical lists
    std::vector<int> start[3], sorted[3];
    // exact linked lists (include/cell.cuh:11-42)
    std::vector<std::array<int, 27>> nbr;
    std::vector<int> head[3], head_prev[3], next[3], next_prev[3];
    bool lists_built = false;

    Oracle(const fsa_params &p, int mode_) : P(convert<T>(p)), mode(mode_) {
        fx.resize(P.nf); fu.resize(P.nf); fa.assign(P.nf, {0, 0, 0});
        fcx.assign(P.nf, {0, 0, 0}); fcu.assign(P.nf, {0, 0, 0});
        scx.assign(P.ns, {0, 0, 0}); scu.assign(P.ns, {0, 0, 0});
        frho.resize(P.nf); fp.resize(P.nf); fdrho.assign(P.nf, T(0));
        sx.resize(P.ns); su.resize(P.ns); sX.resize(P.ns);
        sn.assign(P.ns, {0, 0, 0}); sa.assign(P.ns, {0, 0, 0}); sac.assign(P.ns, {0, 0, 0});
        stype.resize(P.ns); sp.assign(P.ns, T(0));
        sF.assign(P.ns, mident<T>()); sP.assign(P.ns, mzero<T>()); sL.assign(P.ns, mident<T>());
        vx.resize(P.nv); vn.assign(P.nv, {0, 0, 0}); vu.assign(P.nv, {0, 0, 0});
        vrho.assign(P.nv, P.rho0); vp.assign(P.nv, T(0)); vtype.resize(P.nv);
        const int n[3] = {P.nf, P.ns, P.nv};
        for (int k = 0; k < 3; ++k) {
            cell[k].assign(n[k], -1);
            sub[k].assign(n[k], 0);
            next[k].assign(n[k], -1);
            next_prev[k].assign(n[k], -1);
            head[k].assign(P.ncell, -1);
            head_prev[k].assign(P.ncell, -1);
        }
        if (mode == EXACT) create_cell();
    }

    int count(int k) const { return k == KF ? P.nf : (k == KS ? P.ns : P.nv); }
    const std::vector<V3<T>> &pos(int k) const { return k == KF ? fx : (k == KS ? sx : vx); }

    // findPositionCell, src/pair.cu:8-24 (host twin Sim::getCellID, src/sim.cu:238-254)
    int find_cell(V3<T> x) const {
        int nx = (int)((x.x - P.cmin[0]) * P.inv[0]);
        int ny = (int)((x.y - P.cmin[1]) * P.inv[1]);
        int nz = (int)((x.z - P.cmin[2]) * P.inv[2]);
        return nx * P.cn[1] * P.cn[2] + ny * P.cn[2] + nz;
    }
    bool in_grid(V3<T> x) const {
        T v[3] = {x.x, x.y, x.z};
        for (int a = 0; a < 3; ++a) {
            if (!(v[a] >= P.cmin[a])) return false;
            int i = (int)((v[a] - P.cmin[a]) * P.inv[a]);
            if (i < 0 || i >= P.cn[a]) return false;
        }
        return true;
    }

    // Sim::createCell, src/sim.cu:153-236: neighbour table in the fixed order
    void create_cell() {
        static const int off[27][3] = {
            {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {-1, 0, -1}, {-1, 0, 1}, {-1, -1, 0},
            {-1, 1, 0},   {-1, 0, 0},  {0, -1, -1}, {0, -1, 1}, {0, 1, -1},  {0, 1, 1},  {0, 0, -1},
            {0, 0, 1},    {0, -1, 0},  {0, 1, 0},   {0, 0, 0},  {1, -1, -1}, {1, -1, 1}, {1, 1, -1},
            {1, 1, 1},    {1, 0, -1},  {1, 0, 1},   {1, -1, 0}, {1, 1, 0},   {1, 0, 0}};
        nbr.assign(P.ncell, {});
        for (int i = 0; i < P.cn[0]; ++i)
            for (int j = 0; j < P.cn[1]; ++j)
                for (int k = 0; k < P.cn[2]; ++k) {
                    int id = (i * P.cn[1] + j) * P.cn[2] + k;
                    for (int n = 0; n < 27; ++n) {
                        int a = i + off[n][0], b = j + off[n][1], c = k + off[n][2];
                        bool ok = a >= 0 && a < P.cn[0] && b >= 0 && b < P.cn[1] && c >= 0 && c < P.cn[2];
                        nbr[id][n] = ok ? (a * P.cn[1] + b) * P.cn[2] + c : -1;
                    }
                }
    }

    // Sim::fillCell (host), src/sim.cu:256-302: append in ascending index
    void host_fill_cell() {
        for (int k = 0; k < 3; ++k) {
            std::fill(head[k].begin(), head[k].end(), -1);
            std::fill(next[k].begin(), next[k].end(), -1);
            std::vector<int> tail(P.ncell, -1);
            const auto &x = pos(k);
            for (int i = 0; i < count(k); ++i) {
                int c = find_cell(x[i]);
                if (head[k][c] == -1) head[k][c] = i;
                else next[k][tail[c]] = i;
                tail[c] = i;
            }
        }
        lists_built = true;
    }

    // ---- traversal -----------------------------------------------------
    template <class F> void visit(int c, unsigned mask, F &&f) const {
        if (mode == CANONICAL) {
            const int ny = P.cn[1], nz = P.cn[2];
            const int cx = c / (ny * nz), cy = (c / nz) % ny, cz = c % nz;
            const int zlo = std::max(cz - 1, 0), zhi = std::min(cz + 1, nz - 1);
            for (int k = 0; k < 3; ++k) {
                if (!(mask & (1u << k))) continue;
                for (int ox = -1; ox <= 1; ++ox) {
                    const int gx = cx + ox;
                    if (gx < 0 || gx >= P.cn[0]) continue;
                    for (int oy = -1; oy <= 1; ++oy) {
                        const int gy = cy + oy;
                        if (gy < 0 || gy >= ny) continue;
                        const int base = (gx * ny + gy) * nz;
                        const int b = start[k][base + zlo], e = start[k][base + zhi + 1];
                        for (int t = b; t < e; ++t) f(k, sorted[k][t]);
                    }
                }
            }
        } else {
            for (int n = 0; n < 27; ++n) {
                const int nc = nbr[c][n];
                if (nc == -1) continue;
                for (int k = 0; k < 3; ++k) {
                    if (!(mask & (1u << k))) continue;
                    for (int j = head[k][nc]; j != -1; j = next[k][j]) f(k, j);
                }
            }
        }
    }

    // ---- uploads (*_Host_List::addList, include/host_data.cuh:20-163) ----
    int upload_fluid(const float *x, const float *u, const float *rho, const float *p) override {
        for (int i = 0; i < P.nf; ++i) {
            fx[i] = {T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2])};
            fu[i] = {T(u[3 * i]), T(u[3 * i + 1]), T(u[3 * i + 2])};
            frho[i] = rho[i];
            fp[i] = p[i];
            if (!in_grid(fx[i])) { err = "fluid particle outside the cell grid"; return FSA_ERR_INVALID; }
        }
        fcx.assign(P.nf, {0, 0, 0});
        fcu.assign(P.nf, {0, 0, 0});
        lists_built = false;
        return FSA_OK;
    }
    int upload_solid(const float *x, const float *u, const float *X, const int32_t *type) override {
        for (int i = 0; i < P.ns; ++i) {
            sx[i] = {T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2])};
            su[i] = {T(u[3 * i]), T(u[3 * i + 1]), T(u[3 * i + 2])};
            sX[i] = {T(X[3 * i]), T(X[3 * i + 1]), T(X[3 * i + 2])};
            stype[i] = type[i];
            if (!in_grid(sx[i])) { err = "solid particle outside the cell grid"; return FSA_ERR_INVALID; }
        }
        scx.assign(P.ns, {0, 0, 0});
        scu.assign(P.ns, {0, 0, 0});
        lists_built = false;
        return FSA_OK;
    }
    int upload_virt(const float *x, const int32_t *type) override {
        for (int i = 0; i < P.nv; ++i) {
            vx[i] = {T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2])};
            vtype[i] = type[i];
            if (!in_grid(vx[i])) { err = "virtual particle outside the cell grid"; return FSA_ERR_INVALID; }
        }
        lists_built = false;
        return FSA_OK;
    }
    static void put3(double *o, const std::vector<V3<T>> &v) {
        if (!o) return;
        for (size_t i = 0; i < v.size(); ++i) {
            o[3 * i] = v[i].x; o[3 * i + 1] = v[i].y; o[3 * i + 2] = v[i].z;
        }
    }
    static void put1(double *o, const std::vector<T> &v) {
        if (!o) return;
        for (size_t i = 0; i < v.size(); ++i) o[i] = v[i];
    }
    // column-major (glm) 9 doubles per matrix
    static void put9(double *o, const std::vector<M3<T>> &v) {
        if (!o) return;
        for (size_t i = 0; i < v.size(); ++i)
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) o[9 * i + 3 * c + r] = v[i].m[r][c];
    }
    int download_fluid(double *x, double *u, double *rho, double *p, double *drho, double *a) override {
        put3(x, fx); put3(u, fu); put1(rho, frho); put1(p, fp); put1(drho, fdrho); put3(a, fa);
        return FSA_OK;
    }
    int download_solid(double *x, double *u, double *p, double *n, double *a, double *ac, double *F, double *Pk,
                       double *L) override {
        put3(x, sx); put3(u, su); put1(p, sp); put3(n, sn); put3(a, sa); put3(ac, sac);
        put9(F, sF); put9(Pk, sP); put9(L, sL);
        return FSA_OK;
    }
    int download_virt(double *n, double *u, double *rho, double *p) override {
        put3(n, vn); put3(u, vu); put1(rho, vrho); put1(p, vp);
        return FSA_OK;
    }

    // ---- K1-K3: particleCellUpdate, src/pair.cu:26-70 -------------------
    void particle_cell_update() override {
        for (int k = 0; k < 3; ++k) {
            const auto &x = pos(k);
            const int n = count(k);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                cell[k][i] = find_cell(x[i]);
                const int ns = k == KS ? 1 : kSubZ;
                const T tz = (x[i].z - P.cmin[2]) * P.inv[2];
                const int iz = std::min(std::max((int)tz, 0), P.cn[2] - 1);
                sub[k][i] = std::min(std::max((int)(tz * T(ns)) - ns * iz, 0), ns - 1);
            }
        }
    }

    int64_t lost = 0;
    int64_t lost_count() override { return lost; }

    // ---- cell lists ------------------------------------------------------
    void fill_cell() override {
        if (mode == CANONICAL) {
            // stable counting sort by (cell, eighth-cell) over ascending .dat index
            for (int k = 0; k < 3; ++k) {
                const int n = count(k);
                const int ns = k == KS ? 1 : kSubZ;
                const int nk = P.ncell * ns;
                std::vector<int> fstart(nk + 1, 0);
                for (int i = 0; i < n; ++i) fstart[cell[k][i] * ns + sub[k][i] + 1]++;
                for (int c = 0; c < nk; ++c) fstart[c + 1] += fstart[c];
                sorted[k].assign(n, -1);
                std::vector<int> fillp(fstart.begin(), fstart.end() - 1);
                for (int i = 0; i < n; ++i) sorted[k][fillp[cell[k][i] * ns + sub[k][i]]++] = i;
                start[k].assign(P.ncell + 1, 0);
                for (int c = 0; c <= P.ncell; ++c) start[k][c] = fstart[c * ns];
            }
            return;
        }
        if (!lists_built) host_fill_cell();
        // copyCellList, src/pair.cu:152-193
        for (int k = 0; k < 3; ++k) {
            for (int c = 0; c < P.ncell; ++c) {
                for (int j = head[k][c]; j != -1;) {
                    next_prev[k][j] = next[k][j];
                    next[k][j] = -1;
                    j = next_prev[k][j];
                }
                head_prev[k][c] = head[k][c];
                head[k][c] = -1;
            }
        }
        // fillCellList, src/pair.cu:72-151: one writer per cell
        for (int k = 0; k < 3; ++k) {
#pragma omp parallel for schedule(dynamic, 256)
            for (int c = 0; c < P.ncell; ++c) {
                int current = -1;
                for (int n = 0; n < 27; ++n) {
                    const int nc = nbr[c][n];
                    if (nc == -1) continue;
                    for (int j = head_prev[k][nc]; j != -1; j = next_prev[k][j]) {
                        if (cell[k][j] != c) continue;
                        if (head[k][c] == -1) head[k][c] = j;
                        else next[k][current] = j;
                        current = j;
                    }
                }
            }
        }
        // particles present in no list are lost for good (reference behaviour)
        int64_t in_lists = 0;
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < P.ncell; ++c)
                for (int j = head[k][c]; j != -1; j = next[k][j]) ++in_lists;
        lost = (int64_t)P.nf + P.ns + P.nv - in_lists;
    }

    // ---- K6: solidCoupleUpdate, src/solid.cu:262-344 ---------------------
    void couple_calculation() override {
        if (P.nf == 0 || P.ns == 0) return;  // coupleCalculation, src/solid.cu:346-352
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < P.ns; ++i) {
            const V3<T> xi = sx[i];
            V3<T> normal = {0, 0, 0};
            T pressure = 0, sum_correction = 0;
            visit(cell[KS][i], (1u << KF) | (1u << KS), [&](int k, int j) {
                const V3<T> xj = k == KF ? fx[j] : sx[j];
                const V3<T> dx = xj - xi;
                const T d2 = dist2(dx);
                if (!(d2 < P.R2)) return;
                const T q2 = d2 * P.h_inv2;
                if (k == KF) {  // solidPressureCalculate, src/solid.cu:275-291
                    const T w = (std::exp(-q2) - P.c1) * P.c2;
                    sum_correction += w;
                    pressure += fp[j] * w;
                } else {  // solidNormalCalculate, src/solid.cu:262-273
                    const T kc = std::exp(-q2) * P.kd;
                    normal -= dx * kc;
                }
            });
            sn[i] = length2(normal) < T(kEpsForSum) ? V3<T>{0, 0, 0} : normalize(normal);
            sp[i] = sum_correction > T(kEpsFloat) ? std::fmax(pressure / sum_correction, P.p_min) : T(0);
        }
    }

    // ---- K7: getVirtData, src/virt.cu:9-75 --------------------------------
    void virt_update() override {
        if (P.nv == 0) return;
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < P.nv; ++i) {
            const V3<T> xi = vx[i];
            V3<T> velocity = {0, 0, 0}, normal = {0, 0, 0};
            T pressure = 0, density = P.rho0, nearest = T(kLargeFloat);
            visit(cell[KV][i], (1u << KF) | (1u << KV), [&](int k, int j) {
                if (k == KF) {
                    const V3<T> dx = fx[j] - xi;
                    const T d2 = dist2(dx);
                    if (d2 < nearest) {
                        density = frho[j];
                        pressure = fp[j] - density * dot(P.g, dx);
                        velocity = fu[j];
                        nearest = d2;
                    }
                } else {
                    const V3<T> dx = vx[j] - xi;
                    const T d2 = dist2(dx);
                    if (d2 < P.R2 && i != j) {
                        const T q2 = d2 * P.h_inv2;
                        const T kc = std::exp(-q2) * P.kd;
                        normal += (-dx) * kc;
                    }
                }
            });
            normal = length2(normal) < T(kEpsForSum) ? V3<T>{0, 0, 0} : normalize(normal);
            vn[i] = normal;
            vp[i] = pressure;
            vu[i] = velocity - (T(2) * dot(velocity, normal)) * normal;
            vrho[i] = density;
        }
    }

    // ---- K8: findFluidPair + pair functions, src/fluid.cu:26-232 ---------
    // the three pair functions differ in the partner state and in two terms;
    // kind: 0 fluid (fluid.cu:26-69), 1 solid (fluid.cu:71-119), 2 virt (fluid.cu:121-167)
    inline void fluid_pair(int kind, V3<T> dx, T d2, V3<T> ui, T rhoi, T pi, V3<T> uj, T rhoj, T pj, T &drho,
                           V3<T> &a) const {
        const T q2 = d2 * P.h_inv2;
        const T kc = std::exp(-q2) * P.kd;
        const V3<T> e = q2 != T(0) ? normalize(dx) : V3<T>{0, 0, 0};
        T ps = (rhoi * pj + rhoj * pi) / (rhoi + rhoj);
        const T du = dot(ui - uj, e);
        if (du >= T(0)) {
            const T diss = kind == 1 ? P.cs : std::fmin(T(3) * du, P.cs);
            ps += rhoi * rhoj * du / (rhoi + rhoj) * diss;
        }
        if (kind == 2) ps = std::fmax(ps, T(0));
        V3<T> vs = (rhoi * ui + rhoj * uj) / (rhoi + rhoj);
        vs += (pi - pj) / (rhoi + rhoj) / P.cs * e;
        a += (T(-2) * ps / rhoi * dx) * kc;
        drho += T(2) * rhoi * dot(ui - vs, dx) * kc;
        if (!P.viscid) return;
        T vis = dot(uj - ui, dx);
        vis /= length2(dx) + T(0.01f) * P.h2;
        a += (T(10) * P.visc * vis / rhoi * dx) * kc;
    }

    // Tait equation of state, densityToPressure src/fluid.cu:16-24
    T density_to_pressure(T rho) const {
        T p = (std::pow(rho * P.rho0_inv, P.gamma) - T(1)) * P.rho2p;
        return std::fmax(p, P.p_min);
    }

    void fluid_calculation() override {
        if (P.nf == 0) return;
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < P.nf; ++i) {
            const V3<T> xi = fx[i], ui = fu[i];
            const T rhoi = frho[i], pi = fp[i];
            T drho = 0;
            V3<T> a = {0, 0, 0};
            visit(cell[KF][i], 7u, [&](int k, int j) {
                const V3<T> xj = pos(k)[j];
                const V3<T> dx = xj - xi;
                const T d2 = dist2(dx);
                if (!(d2 < P.R2)) return;
                if (k == KF) {
                    fluid_pair(0, dx, d2, ui, rhoi, pi, fu[j], frho[j], fp[j], drho, a);
                } else if (k == KS) {
                    // mirrored solid velocity, fluid.cu:79-83; rho_j = rho_i, p_j = p_s
                    const V3<T> n = sn[j];
                    const V3<T> uj = ui - (T(2) * dot(ui - su[j], n)) * n;
                    fluid_pair(1, dx, d2, ui, rhoi, pi, uj, rhoi, sp[j], drho, a);
                } else {
                    fluid_pair(2, dx, d2, ui, rhoi, pi, vu[j], vrho[j], vp[j], drho, a);
                }
            });
            fdrho[i] = drho;
            fa[i] = a;
        }
        // K9: fluidUpdate, src/fluid.cu:234-269
        const T max_speed = T(0.3f) * P.cs;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < P.nf; ++i) {
            if (!(fdrho[i] == fdrho[i])) fdrho[i] = 0;
            const V3<T> a = fa[i];
            if (!(a.x == a.x && a.y == a.y && a.z == a.z)) fa[i] = {0, 0, 0};
            if (P.compensated) {
                step_compensated(fx[i], fu[i] * P.dt, fcx[i]);
                step_compensated(fu[i], (fa[i] + P.g) * P.dt, fcu[i]);
            } else {
                fx[i] += fu[i] * P.dt;
                fu[i] += (fa[i] + P.g) * P.dt;
            }
            frho[i] += fdrho[i] * P.dt;
            frho[i] = std::fmin(std::fmax(frho[i], P.rho_min), P.rho_max);
            fp[i] = density_to_pressure(frho[i]);
            clamp_to_domain(fx[i], fcx[i]);
            const T speed = std::sqrt(length2(fu[i]));
            if (speed > max_speed) {
                fu[i] = fu[i] * (max_speed / speed);
                fcu[i] = {0, 0, 0};  // a rescaled velocity drops its carried error
            }
        }
    }

    // ---- K10: solidDeformationCalculate, src/solid.cu:10-79 --------------
    void solid_deformation() override {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < P.ns; ++i) {
            M3<T> correction = mzero<T>(), deformation = mzero<T>();
            const V3<T> Xi = sX[i];
            visit(cell[KS][i], 1u << KS, [&](int, int j) {
                const V3<T> dX = sX[j] - Xi;
                if (!(dist2(dX) < P.R2)) return;
                const V3<T> dx = sx[j] - sx[i];
                const T q2 = length2(dX) * P.h_inv2;
                const T kc = std::exp(-q2) * P.kd;
                deformation = deformation + outer(dx, dX) * kc;
                correction = correction + outer(dX, dX) * kc;
            });
            correction = det(correction) != T(0) ? inverse(correction) : mident<T>();
            deformation = deformation * correction;
            sF[i] = deformation;
            // solidStressCalculation, src/solid.cu:27-36 (St Venant-Kirchhoff)
            const M3<T> green = T(0.5f) * (transpose(deformation) * deformation - mident<T>());
            const T trace = green.m[0][0] + green.m[1][1] + green.m[2][2];
            const M3<T> pk2 = T(2) * P.mu * green + P.lambda * trace * mident<T>();
            sP[i] = deformation * pk2;
            sL[i] = correction;
        }
    }

    // ---- K11: solidForceCalculate, src/solid.cu:81-207 -------------------
    void solid_force() override {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < P.ns; ++i) {
            V3<T> couple = {0, 0, 0}, acc = {0, 0, 0};
            const V3<T> xi = sx[i], Xi = sX[i];
            const M3<T> PLi = sP[i] * sL[i];
            const M3<T> FinvT = transpose(inverse(sF[i]));
            const T detF = det(sF[i]);
            visit(cell[KS][i], (1u << KF) | (1u << KS), [&](int k, int j) {
                if (k == KF) {
                    const V3<T> dx = fx[j] - xi;
                    if (!(dist2(dx) < P.R2)) return;
                    // solidFluidCouple, src/solid.cu:81-117
                    const V3<T> n = sn[i];
                    const V3<T> uj = fu[j];
                    const V3<T> ui = uj - (T(2) * dot(uj - su[i], n)) * n;
                    const T q2 = length2(dx) * P.h_inv2;
                    const T kc = std::exp(-q2) * P.kd;
                    const T p_i = sp[i], p_j = fp[j];
                    const T rhoj = frho[j], rhoi = rhoj;
                    const V3<T> e = std::sqrt(length2(dx)) != T(0) ? normalize(dx) : V3<T>{0, 0, 0};
                    T ps = (rhoi * p_j + rhoj * p_i) / (rhoi + rhoj);
                    const T du = dot(ui - uj, e);
                    if (du >= T(0)) ps += rhoi * rhoj * du / (rhoi + rhoj) * P.cs;
                    couple += (T(-2) * ps / P.rho_s * dx) * kc;
                } else {
                    const V3<T> dX = sX[j] - Xi;
                    if (!(dist2(dX) < P.R2)) return;
                    // solidSolidForce, src/solid.cu:119-157
                    const T q2 = length2(dX) * P.h_inv2;
                    const T kc = std::exp(-q2) * P.kd;
                    acc += (((sP[j] * sL[j] + PLi) * dX) * kc) / P.rho_s;
                    const V3<T> dx = sx[j] - xi;
                    const V3<T> du = su[j] - su[i];
                    const T du_dot_dx = dot(du, dx);
                    const T eps = T(0.01f) * P.h2;
                    const T nu = P.h * du_dot_dx / (length2(dx) + eps);
                    T viscosity = P.alpha * nu * P.cs_s;
                    viscosity -= P.beta * nu * nu;
                    acc += ((((detF * FinvT) * viscosity) * sL[i]) * dX) * kc;
                }
            });
            sac[i] = couple;
            sa[i] = acc + couple;
        }
    }

    // x += d per component, compensated
    static void step_compensated(V3<T> &x, V3<T> d, V3<T> &c) {
        x.x = kahan_add(x.x, d.x, c.x);
        x.y = kahan_add(x.y, d.y, c.y);
        x.z = kahan_add(x.z, d.z, c.z);
    }
    // the domain clamp of src/fluid.cu:258-260 and src/solid.cu:229-231; a
    // clamped component drops its carried error
    void clamp_to_domain(V3<T> &x, V3<T> &c) const {
        T *v[3] = {&x.x, &x.y, &x.z};
        T *e[3] = {&c.x, &c.y, &c.z};
        for (int a = 0; a < 3; ++a) {
            const T q = std::fmin(std::fmax(*v[a], P.dmin[a]), P.dmax[a]);
            if (q != *v[a]) *e[a] = 0;
            *v[a] = q;
        }
    }

    // ---- K12: solidUpdate, src/solid.cu:209-239 --------------------------
    void solid_update() override {
        const T dt = P.dt / T(P.sub);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < P.ns; ++i) {
            if (stype[i] == FSA_FIXED_SOLID) {
                su[i] = {0, 0, 0};
                sx[i] += su[i] * dt;
                sa[i] = {0, 0, 0};
            } else if (stype[i] == FSA_UNCONSTRAINED_SOLID) {
                if (P.compensated) {
                    step_compensated(sx[i], su[i] * dt, scx[i]);
                    step_compensated(su[i], (sa[i] + P.g) * dt, scu[i]);
                } else {
                    sx[i] += su[i] * dt;
                    su[i] += (sa[i] + P.g) * dt;
                }
                clamp_to_domain(sx[i], scx[i]);
            }
        }
    }

    // solidCalculation, src/solid.cu:241-260
    void solid_calculation() override {
        if (P.ns == 0) return;
        for (int s = 0; s < P.sub; ++s) {
            solid_deformation();
            solid_force();
            solid_update();
        }
    }

    // ---- introspection ---------------------------------------------------
    int cell_ids(int k, int32_t *out) override {
        for (int i = 0; i < count(k); ++i) out[i] = cell[k][i];
        return FSA_OK;
    }
    int cell_start(int k, int32_t *out) override {
        if (mode != CANONICAL) { err = "cell_start needs the canonical mode"; return FSA_ERR_STATE; }
        for (int c = 0; c <= P.ncell; ++c) out[c] = start[k][c];
        return FSA_OK;
    }
    int sorted_ids(int k, int32_t *out) override {
        if (mode != CANONICAL) { err = "sorted_ids needs the canonical mode"; return FSA_ERR_STATE; }
        for (int i = 0; i < count(k); ++i) out[i] = sorted[k][i];
        return FSA_OK;
    }
    V3<T> metric_pos(int k, int j, int ref) const { return (k == KS && ref) ? sX[j] : pos(k)[j]; }
    int neighbors(int ki, int i, unsigned mask, int ref, int32_t *ok, int32_t *oj, int cap) override {
        int n = 0;
        const V3<T> xi = metric_pos(ki, i, ref);
        visit(cell[ki][i], mask, [&](int k, int j) {
            const V3<T> d = metric_pos(k, j, ref) - xi;
            if (dist2(d) < P.R2) {
                if (n < cap) { ok[n] = k; oj[n] = j; }
                ++n;
            }
        });
        return n;
    }
    // findFluidPair's neighbour sets (src/fluid.cu:188-227) as counts: per
    // fluid particle, the fluid, solid and virtual particles within the
    // cut-off (itself included), from the 27-cell walk or from every particle
    void pair_counts(int brute, int32_t *out) override {
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < P.nf; ++i) {
            int32_t c[3] = {0, 0, 0};
            const V3<T> xi = fx[i];
            if (brute) {
                for (int k = 0; k < 3; ++k)
                    for (int j = 0; j < count(k); ++j)
                        if (dist2(metric_pos(k, j, 0) - xi) < P.R2) ++c[k];
            } else {
                visit(cell[KF][i], 7u, [&](int k, int j) {
                    if (dist2(metric_pos(k, j, 0) - xi) < P.R2) ++c[k];
                });
            }
            for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = c[k];
        }
    }
    int brute_neighbors(int ki, int i, unsigned mask, int ref, int32_t *ok, int32_t *oj, int cap) override {
        int n = 0;
        const V3<T> xi = metric_pos(ki, i, ref);
        for (int k = 0; k < 3; ++k) {
            if (!(mask & (1u << k))) continue;
            for (int j = 0; j < count(k); ++j) {
                const V3<T> d = metric_pos(k, j, ref) - xi;
                if (dist2(d) < P.R2) {
                    if (n < cap) { ok[n] = k; oj[n] = j; }
                    ++n;
                }
            }
        }
        return n;
    }
};

thread_local std::string g_err;

}  // namespace

struct orc_sim {
    OracleBase *impl;
};

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int orc_create(const fsa_params *p, int precision, int mode, orc_sim **out) {
    if (!p || !out || (precision != 32 && precision != 64) || (mode != CANONICAL && mode != EXACT)) {
        g_err = "orc_create: bad argument";
        return FSA_ERR_INVALID;
    }
    if (p->cell_number_total <= 0 || p->n_fluid < 0 || p->n_solid < 0 || p->n_virt < 0) {
        g_err = "orc_create: bad parameter block";
        return FSA_ERR_INVALID;
    }
    orc_sim *s = new orc_sim;
    if (precision == 32) s->impl = new Oracle<float>(*p, mode);
    else s->impl = new Oracle<double>(*p, mode);
    *out = s;
    return FSA_OK;
}

int orc_destroy(orc_sim *s) {
    if (s) {
        delete s->impl;
        delete s;
    }
    return FSA_OK;
}

#define ORC_CALL(expr)                              \
    do {                                            \
        int rc_ = (expr);                           \
        if (rc_ != FSA_OK) g_err = s->impl->err;    \
        return rc_;                                 \
    } while (0)

int orc_upload_fluid(orc_sim *s, const float *x, const float *u, const float *rho, const float *p) {
    ORC_CALL(s->impl->upload_fluid(x, u, rho, p));
}
int orc_upload_solid(orc_sim *s, const float *x, const float *u, const float *X, const int32_t *t) {
    ORC_CALL(s->impl->upload_solid(x, u, X, t));
}
int orc_upload_virt(orc_sim *s, const float *x, const int32_t *t) { ORC_CALL(s->impl->upload_virt(x, t)); }
int orc_download_fluid(orc_sim *s, double *x, double *u, double *rho, double *p, double *drho, double *a) {
    ORC_CALL(s->impl->download_fluid(x, u, rho, p, drho, a));
}
int orc_download_solid(orc_sim *s, double *x, double *u, double *p, double *n, double *a, double *ac, double *F,
                       double *Pk, double *L) {
    ORC_CALL(s->impl->download_solid(x, u, p, n, a, ac, F, Pk, L));
}
int orc_download_virt(orc_sim *s, double *n, double *u, double *rho, double *p) {
    ORC_CALL(s->impl->download_virt(n, u, rho, p));
}
int orc_particle_cell_update(orc_sim *s) { s->impl->particle_cell_update(); return FSA_OK; }
int orc_fill_cell(orc_sim *s) { s->impl->fill_cell(); return FSA_OK; }
int orc_couple_calculation(orc_sim *s) { s->impl->couple_calculation(); return FSA_OK; }
int orc_virt_update(orc_sim *s) { s->impl->virt_update(); return FSA_OK; }
int orc_fluid_calculation(orc_sim *s) { s->impl->fluid_calculation(); return FSA_OK; }
int orc_solid_calculation(orc_sim *s) { s->impl->solid_calculation(); return FSA_OK; }
int orc_pair_counts(orc_sim *s, int brute, int32_t *out) { s->impl->pair_counts(brute, out); return FSA_OK; }
int orc_solid_deformation(orc_sim *s) { s->impl->solid_deformation(); return FSA_OK; }
int orc_solid_force(orc_sim *s) { s->impl->solid_force(); return FSA_OK; }
int orc_solid_update(orc_sim *s) { s->impl->solid_update(); return FSA_OK; }
int orc_step(orc_sim *s, int n) {
    for (int i = 0; i < n; ++i) s->impl->step();
    return FSA_OK;
}
int orc_get_cell_ids(orc_sim *s, int kind, int32_t *out) { ORC_CALL(s->impl->cell_ids(kind, out)); }
int orc_get_cell_start(orc_sim *s, int kind, int32_t *out) { ORC_CALL(s->impl->cell_start(kind, out)); }
int orc_get_sorted_ids(orc_sim *s, int kind, int32_t *out) { ORC_CALL(s->impl->sorted_ids(kind, out)); }
int orc_neighbors(orc_sim *s, int kind_i, int i, unsigned mask, int reference_metric, int32_t *ok, int32_t *oj,
                  int cap) {
    return s->impl->neighbors(kind_i, i, mask, reference_metric, ok, oj, cap);
}
int orc_brute_neighbors(orc_sim *s, int kind_i, int i, unsigned mask, int reference_metric, int32_t *ok,
                        int32_t *oj, int cap) {
    return s->impl->brute_neighbors(kind_i, i, mask, reference_metric, ok, oj, cap);
}
int64_t orc_lost_count(orc_sim *s) { return s->impl->lost_count(); }

}  // extern "C"
