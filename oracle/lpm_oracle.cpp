// FP64 CPU oracle of the LPM-MAC static / fatigue Newton load-step path.
//
// TEST INFRASTRUCTURE ONLY. This file restates the reference algorithm so that
// the HIP implementation can be checked against it. Only tests/, the smoke()
// check in __graft_entry__.py and the cpu_baseline leg of bench.py may load it;
// the product library never links, loads or calls it.
//
// Every function cites the reference file:line it restates (paths are relative
// to the reference repository root). Floating-point operations are written in
// the same order as the reference so that FP64 results agree to the last bit
// where the reference's order is known. Where the reference delegates to Eigen
// 3.4 (CG, dot products) the Eigen algorithm is restated from its published
// source; the summation order inside Eigen's vectorised dot product is not
// restated (sequential sums are used), so CG iterates can differ from Eigen at
// the ulp level: parity of the CG inner products is unpinned at that level.
//
// Build: see oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off, no -march,
// so no FMA is formed, matching a default x86-64 build of the reference).

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

namespace ora {

constexpr double PI = 3.14159265358979323846;  // include/lpm.h:36
constexpr double EPS = 1e-6;                   // include/lpm.h:37
constexpr int NDIM = 3;                        // include/lpm.h:38
constexpr int NL = 2;                          // nlayer; the reference supports 2 layers only
constexpr int NSV = 7;                         // max state_var length (J2: 7)

enum Lattice { Square2D = 0, Hexagon2D, SimpleCubic3D, FCC3D, BCC3D_1Slip, BCC3D_2Slip, BCC3D_3Slip };
enum PType { Elastic = 0, ElasticDamage, J2Plasticity, FatigueHCF };
enum LoadModeE { Relative = 0, Absolute = 1 };

// func_phi: include/lpm.h:45-49
static double func_phi(const double x, const double L) {
    double p = 8, q = 2;
    return std::pow(1 / (1 + std::pow(x / L, p)), q);
}

// UnitCell: include/unit_cell.h:7-127
struct Cell {
    int lattice = 0, dim = 0;
    int nneighbors = 0, nneighbors1 = 0, nneighbors2 = 0;
    int nAFEM = 0, nAFEM1 = 0, nAFEM2 = 0;
    double radius = 0, volume = 0;
    double cutoff[2] = {0, 0};
    double data[9] = {0};  // el_mapping data array, in the order written in the source
};

static Cell make_cell(int lattice, double r) {
    Cell c;
    c.lattice = lattice;
    c.radius = r;
    if (lattice == Square2D) {  // unit_cell.h:44-59
        double d[9] = {1. / 2., -1. / 2., 0., 0., 0., 1. / 2., 0., 1. / 12., -1. / 12.};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 2; c.nneighbors = 8; c.nneighbors1 = 4; c.nneighbors2 = 4;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 2.0 * std::sqrt(2.0) * r;
        c.nAFEM = 16; c.nAFEM1 = 12; c.nAFEM2 = 12;
        c.volume = 4 * std::pow(r, 2);
    } else if (lattice == Hexagon2D) {  // unit_cell.h:61-76
        double d[9] = {std::sqrt(3.0) / 12.0, -std::sqrt(3.0) / 12.0, 0.,
                       -std::sqrt(3.0) / 144.0, std::sqrt(3.0) / 48.0, 0.0,
                       0.0, 0.0, 1.0};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 2; c.nneighbors = 12; c.nneighbors1 = 6; c.nneighbors2 = 6;
        c.nAFEM = 30; c.nAFEM1 = 18; c.nAFEM2 = 18;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 2.0 * std::sqrt(3.0) * r;
        c.volume = 2 * std::sqrt(3) * std::pow(r, 2);
    } else if (lattice == SimpleCubic3D) {  // unit_cell.h:78-93
        double d[9] = {1, -1, -1.0, 0.0, 0.0, 1., 0.0, 1 / 18.0, -1 / 18.0};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 3; c.nneighbors = 18; c.nneighbors1 = 6; c.nneighbors2 = 12;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 2.0 * std::sqrt(2.0) * r;
        c.nAFEM = 60; c.nAFEM1 = 24; c.nAFEM2 = 54;
        c.volume = std::pow(2 * r, 3);
    } else if (lattice == FCC3D) {  // unit_cell.h:95-110
        double d[9] = {0.0, 0.0, std::sqrt(2.0), std::sqrt(2.0) / 4, -std::sqrt(2.0) / 4, -std::sqrt(2.0) / 4,
                       0.0, std::sqrt(2.0) / 24, -std::sqrt(2.0) / 24};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 3; c.nneighbors = 18; c.nneighbors1 = 12; c.nneighbors2 = 6;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 2.0 * std::sqrt(2.0) * r;
        c.nAFEM = 60; c.nAFEM1 = 54; c.nAFEM2 = 24;
        c.volume = 4.0 * std::sqrt(2.0) * std::pow(r, 3);
    } else {  // BCC, unit_cell.h:112-127
        double d[9] = {0.0, 0.0, std::sqrt(3.0), 1 / std::sqrt(3.0), -1 / std::sqrt(3.0), 0.,
                       0.0, std::sqrt(3.0) / 14, std::sqrt(3.0) / 14};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 3; c.nneighbors = 14; c.nneighbors1 = 8; c.nneighbors2 = 6;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 4.0 / std::sqrt(3.0) * r;
        c.nAFEM = 40; c.nAFEM1 = 34; c.nAFEM2 = 24;
        c.volume = 32.0 * std::sqrt(3.0) / 9.0 * std::pow(r, 3);
    }
    return c;
}

// el_mapping(r, c). The reference builds it with Eigen::Map, which is
// column-major, from a row-major-written array (unit_cell.h:49,66,83,100,117):
// literal => M(r,c) = data[c*3 + r]; intended (row-major) => data[r*3 + c].
static double el_map(const Cell &c, bool literal, int r, int col) {
    return literal ? c.data[col * 3 + r] : c.data[r * 3 + col];
}

struct LoadStep {
    struct Disp { std::vector<int> group; int mode; char flag; double step; };
    struct Force { std::vector<int> group; int mode; double fx, fy, fz; };
    std::vector<Disp> disp;
    std::vector<Force> force;
    // LoadStep::loadCutHalf, include/load_step.h:21-28; boundary.h:21-24,40-43
    void cutHalf() {
        for (auto &d : disp) d.step /= 2.0;
        for (auto &f : force) { f.fx /= 2.0; f.fy /= 2.0; f.fz /= 2.0; }
    }
};

struct Sys {
    Cell cell;
    int ptype = Elastic;
    int N = 0, dim = 0;
    double box[6] = {0};
    // mode switches (each defaults to the reference's literal behaviour)
    int elmap_literal = 1;   // column-major read of el_mapping
    int diag10_literal = 1;  // diagonal block K[1][0] left at 0
    int cg_transpose = 1;    // Eigen CG with Lower|Upper on column-major K solves K^T d = r
    int dirichlet_static = 0;// static solver does not eliminate Dirichlet DoFs
    int rollback_pex = 0;    // Pex is not rolled back on a step restart
    int stiffness_fd = 0;    // StiffnessMode::FiniteDifference
    int fd_exact = 0;        // FD with exact restores and a fresh base force (not the reference)
    int cg_forcing = 0;      // inexact Newton: per-iteration CG tolerance (not the reference)
    double cg_tol = 2.220446049250313e-16;  // Eigen default tolerance (epsilon)
    long cg_maxit = -1;                     // -1 => 2 * ndof (Eigen default)
    int max_NR_iter = 30;
    double tol_NR = 1e-6;
    int verbose = 0;

    // particles
    std::vector<double> xyz, xyz0, xyz_last;  // 3N
    std::vector<int> type;
    std::vector<int> nb;                      // number of bonds (= neighbors.size())
    std::vector<double> dLe_total, TdLe_total; // 2N
    std::vector<double> cs_sum;               // 6N: [i][layer][3]
    std::vector<double> Pin, Pex;             // 3N
    std::vector<int> cons;                    // 3N disp_constraint
    std::vector<double> damage, damage_last, damage_visual, Ddot_local, Ddot_nonlocal, nonlocal_L;
    std::vector<double> stress;               // 6N
    std::vector<double> sv, sv_last;          // NSV*N
    int nsv = 1;
    std::vector<double> mat;                  // 10 per particle, model-specific
    // bonds (CSR by particle; layer-0 bonds first, each layer in creation order)
    std::vector<int> bptr, bsplit, p2, layer, opp;
    std::vector<double> dis0, dis, dis_last, dL, dLe, dLp, dLp_last, bstrain, Kn, Tv, cs, bforce, bforce_last,
        bdamage, bdamage_last;
    // 2-hop connectivity (sparsity pattern), assembly.h:236-266
    std::vector<int> cptr, cidx, nconn_largeq;
    // nonlocal neighbours, assembly.h:117-130
    std::vector<int> nlptr, nlidx;
    // stiffness matrix K as CSR over K's rows (row = i*dim+r, col = j*dim+s)
    std::vector<int> krow, kcol;
    std::vector<double> kval;
    std::vector<double> residual, disp, reaction;
    int cg_iters = 0; double cg_error = 0; int cg_success = 0;
    int last_newton = 0;
    int n_steps_done = 0;
    int n_restarts = 0;
    std::vector<int> newton_log;  // Newton iteration count per solveProblemStep call (static) / per substep
    std::vector<int> cg_log;      // CG iteration count per linear solve
};

// Particle::distanceTo, include/particle.h:276-280
static inline double distanceTo(const double *a, const double *b) {
    return std::sqrt(std::pow((a[0] - b[0]), 2) + std::pow((a[1] - b[1]), 2) + std::pow((a[2] - b[2]), 2));
}

static void init_particles(Sys &s, int lattice, double radius, int ptype, int n, const double *xyz, const double *box) {
    s.cell = make_cell(lattice, radius);
    s.ptype = ptype;
    s.N = n;
    s.dim = s.cell.dim;
    if (box) std::memcpy(s.box, box, sizeof s.box);
    s.xyz.assign(xyz, xyz + 3 * n);
    s.xyz0 = s.xyz;
    s.xyz_last = s.xyz;
    s.type.assign(n, 0);
    s.nb.assign(n, 0);
    s.dLe_total.assign(2 * n, 0); s.TdLe_total.assign(2 * n, 0);
    s.cs_sum.assign(6 * n, 0);
    s.Pin.assign(3 * n, 0); s.Pex.assign(3 * n, 0); s.cons.assign(3 * n, 0);
    s.damage.assign(n, 0); s.damage_last.assign(n, 0); s.damage_visual.assign(n, 0);
    s.Ddot_local.assign(n, 0); s.Ddot_nonlocal.assign(n, 0); s.nonlocal_L.assign(n, 0);
    s.stress.assign(6 * n, 0);
    s.sv.assign(NSV * n, 0); s.sv_last.assign(NSV * n, 0);
    s.mat.assign(10 * n, 0);
    // state_var initial values: particle_elastic.h:23-29, particle_elastic_damage.h:177-184,
    // particle_j2plasticity.h:26-33, particle_fatigue_hcf.h:27-28,33-34 (INT_MAX stored as double)
    if (ptype == Elastic || ptype == ElasticDamage) s.nsv = 1;
    else if (ptype == J2Plasticity) s.nsv = 7;
    else { s.nsv = 3; }
    if (ptype == FatigueHCF) {
        for (int i = 0; i < n; ++i) {
            s.sv[NSV * i + 1] = (double)INT_MAX;
            s.sv_last[NSV * i + 1] = (double)INT_MAX;
        }
    }
    // ParticleFatigueHCF default damage_threshold{1}, fatigue_limit_ratio{1.108} (particle_fatigue_hcf.h:24-25)
    if (ptype == FatigueHCF)
        for (int i = 0; i < n; ++i) { s.mat[10 * i + 5] = 1.0; s.mat[10 * i + 6] = 1.108; }
    s.residual.assign(s.dim * n, 0);
    s.disp.assign(s.dim * n, 0);
}

static void alloc_bond_arrays(Sys &s) {
    size_t nbd = s.p2.size();
    s.dis0.assign(nbd, 0); s.dis.assign(nbd, 0); s.dis_last.assign(nbd, 0);
    s.dL.assign(nbd, 0); s.dLe.assign(nbd, 0); s.dLp.assign(nbd, 0); s.dLp_last.assign(nbd, 0);
    s.bstrain.assign(nbd, 0); s.Kn.assign(nbd, 0); s.Tv.assign(nbd, 0); s.cs.assign(3 * nbd, 0);
    s.bforce.assign(nbd, 0); s.bforce_last.assign(nbd, 0); s.bdamage.assign(nbd, 0); s.bdamage_last.assign(nbd, 0);
}

// Bond(p1, p2, layer, dis) constructor, include/bond.h:61-73
static void bond_init(Sys &s, int b, int i, int j, double distance) {
    s.dis0[b] = distance; s.dis[b] = distance; s.dis_last[b] = distance;
    const double *x1 = &s.xyz[3 * i], *x2 = &s.xyz[3 * j];
    s.cs[3 * b + 0] = (x1[0] - x2[0]) / s.dis[b];
    s.cs[3 * b + 1] = (x1[1] - x2[1]) / s.dis[b];
    s.cs[3 * b + 2] = (x1[2] - x2[2]) / s.dis[b];
}

// Reverse-bond index. The reference searches p2's bond list of the same layer for the
// first bond pointing back (particle.h:207-209, 229-231); a missing reverse bond is
// undefined behaviour there and an error here.
static int build_opp(Sys &s) {
    size_t nbd = s.p2.size();
    s.opp.assign(nbd, -1);
    for (int i = 0; i < s.N; ++i) {
        for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b) {
            int j = s.p2[b], l = s.layer[b];
            int lo = (l == 0) ? s.bptr[j] : s.bsplit[j];
            int hi = (l == 0) ? s.bsplit[j] : s.bptr[j + 1];
            for (int c = lo; c < hi; ++c)
                if (s.p2[c] == i) { s.opp[b] = c; break; }
            if (s.opp[b] < 0) return -1;
        }
    }
    return 0;
}

// Assembly::createBonds, include/assembly.h:212-234 (all-pairs, p2 in id order).
// A uniform cell grid restricts the candidate set; candidates are visited in id order
// and accepted by the reference's own test, so the lists are identical.
static void create_bonds(Sys &s) {
    const int n = s.N;
    const double c1 = 1.01 * s.cell.cutoff[1], c0 = 1.01 * s.cell.cutoff[0];
    std::vector<std::vector<int>> l0(n), l1(n);
    std::vector<std::vector<double>> d0(n), d1(n);
    // cell grid with cell size >= c1
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], s.xyz[3 * i + k]); hi[k] = std::max(hi[k], s.xyz[3 * i + k]); }
    double h = c1 * 1.0001;
    long nc[3];
    for (int k = 0; k < 3; ++k) nc[k] = std::max(1L, (long)std::floor((hi[k] - lo[k]) / h) + 1);
    auto cell_of = [&](const double *x, long *c) {
        for (int k = 0; k < 3; ++k) { c[k] = (long)std::floor((x[k] - lo[k]) / h); c[k] = std::min(std::max(c[k], 0L), nc[k] - 1); }
    };
    long ncell = nc[0] * nc[1] * nc[2];
    std::vector<int> head(ncell + 1, 0), items(n);
    std::vector<long> pc(n);
    for (int i = 0; i < n; ++i) { long c[3]; cell_of(&s.xyz[3 * i], c); pc[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0]; head[pc[i] + 1]++; }
    for (long c = 0; c < ncell; ++c) head[c + 1] += head[c];
    { std::vector<int> fill(head.begin(), head.end() - 1); for (int i = 0; i < n; ++i) items[fill[pc[i]]++] = i; }
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        long c[3]; cell_of(&s.xyz[3 * i], c);
        std::vector<int> cand;
        for (long dz = -1; dz <= 1; ++dz) for (long dy = -1; dy <= 1; ++dy) for (long dx = -1; dx <= 1; ++dx) {
            long x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
            if (x < 0 || y < 0 || z < 0 || x >= nc[0] || y >= nc[1] || z >= nc[2]) continue;
            long cc = (z * nc[1] + y) * nc[0] + x;
            for (int t = head[cc]; t < head[cc + 1]; ++t) cand.push_back(items[t]);
        }
        std::sort(cand.begin(), cand.end());
        for (int j : cand) {
            double distance = distanceTo(&s.xyz[3 * i], &s.xyz[3 * j]);
            if ((distance < c1) && (i != j)) {
                if (distance < c0) { l0[i].push_back(j); d0[i].push_back(distance); }
                else { l1[i].push_back(j); d1[i].push_back(distance); }
            }
        }
    }
    s.bptr.assign(n + 1, 0); s.bsplit.assign(n, 0);
    for (int i = 0; i < n; ++i) s.bptr[i + 1] = s.bptr[i] + (int)(l0[i].size() + l1[i].size());
    s.p2.assign(s.bptr[n], 0); s.layer.assign(s.bptr[n], 0);
    alloc_bond_arrays(s);
    for (int i = 0; i < n; ++i) {
        int b = s.bptr[i];
        for (size_t k = 0; k < l0[i].size(); ++k, ++b) { s.p2[b] = l0[i][k]; s.layer[b] = 0; bond_init(s, b, i, l0[i][k], d0[i][k]); }
        s.bsplit[i] = b;
        for (size_t k = 0; k < l1[i].size(); ++k, ++b) { s.p2[b] = l1[i][k]; s.layer[b] = 1; bond_init(s, b, i, l1[i][k], d1[i][k]); }
        s.nb[i] = s.bptr[i + 1] - s.bptr[i];
    }
}

// ---- geometry tools (geometry/geo1_CT_configuration.cpp; the dogbone tools repeat these)
// inCircle (:45-49)
static bool in_circle(const double *pt, double cx, double cy, double r) {
    double dist_sq = std::pow(pt[0] - cx, 2.0) + std::pow(pt[1] - cy, 2.0);
    return dist_sq <= r * r;
}
// inNotch (:52-62)
static bool in_notch(const double *pt, double tx, double ty, double hf_width) {
    if (std::abs(pt[1] - ty) <= hf_width) {
        if ((pt[1] - ty) >= 1. / std::sqrt(3.) * (tx - pt[0]) && (pt[1] - ty) <= 1. / std::sqrt(3.) * (-tx + pt[0]))
            return true;
    }
    return false;
}
// inReverseNotch (:64-74)
static bool in_reverse_notch(const double *pt, double tx, double ty, double hf_width) {
    if (std::abs(pt[1] - ty) <= hf_width) {
        if ((pt[1] - ty) <= 1. / std::sqrt(3.) * (tx - pt[0]) && (pt[1] - ty) >= 1. / std::sqrt(3.) * (-tx + pt[0]))
            return true;
    }
    return false;
}
// isValid (:137-171) with the tool's hard-coded cuts as a list: kind 0 circle (x, y, r),
// 1 notch (tip x, tip y, half width), 2 reverse notch
static bool is_valid(const double *pt, int ncut, const int *kind, const double *par) {
    for (int k = 0; k < ncut; ++k) {
        const double *q = &par[3 * k];
        if (kind[k] == 0 && in_circle(pt, q[0], q[1], q[2])) return false;
        if (kind[k] == 1 && in_notch(pt, q[0], q[1], q[2])) return false;
        if (kind[k] == 2 && in_reverse_notch(pt, q[0], q[1], q[2])) return false;
    }
    return true;
}

// Assembly::createBonds restated literally (all pairs): used to pin the cell-grid version.
static void create_bonds_bruteforce(Sys &s) {
    const int n = s.N;
    const double c1 = 1.01 * s.cell.cutoff[1], c0 = 1.01 * s.cell.cutoff[0];
    std::vector<std::vector<int>> l0(n), l1(n);
    std::vector<std::vector<double>> d0(n), d1(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double distance = distanceTo(&s.xyz[3 * i], &s.xyz[3 * j]);
            if ((distance < c1) && (i != j)) {
                if (distance < c0) { l0[i].push_back(j); d0[i].push_back(distance); }
                else { l1[i].push_back(j); d1[i].push_back(distance); }
            }
        }
    s.bptr.assign(n + 1, 0); s.bsplit.assign(n, 0);
    for (int i = 0; i < n; ++i) s.bptr[i + 1] = s.bptr[i] + (int)(l0[i].size() + l1[i].size());
    s.p2.assign(s.bptr[n], 0); s.layer.assign(s.bptr[n], 0);
    alloc_bond_arrays(s);
    for (int i = 0; i < n; ++i) {
        int b = s.bptr[i];
        for (size_t k = 0; k < l0[i].size(); ++k, ++b) { s.p2[b] = l0[i][k]; s.layer[b] = 0; bond_init(s, b, i, l0[i][k], d0[i][k]); }
        s.bsplit[i] = b;
        for (size_t k = 0; k < l1[i].size(); ++k, ++b) { s.p2[b] = l1[i][k]; s.layer[b] = 1; bond_init(s, b, i, l1[i][k], d1[i][k]); }
        s.nb[i] = s.bptr[i + 1] - s.bptr[i];
    }
}

// Assembly::updateConnections, include/assembly.h:236-266
static void update_connections(Sys &s) {
    const int n = s.N;
    std::vector<std::vector<int>> conns(n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        std::vector<int> &c = conns[i];
        for (int l = 0; l < NL; ++l) {
            int lo = l == 0 ? s.bptr[i] : s.bsplit[i], hi = l == 0 ? s.bsplit[i] : s.bptr[i + 1];
            for (int b = lo; b < hi; ++b) {
                int k = s.p2[b];
                c.push_back(k);
                int lo2 = l == 0 ? s.bptr[k] : s.bsplit[k], hi2 = l == 0 ? s.bsplit[k] : s.bptr[k + 1];
                for (int b2 = lo2; b2 < hi2; ++b2) c.push_back(s.p2[b2]);
            }
        }
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
    }
    s.cptr.assign(n + 1, 0); s.nconn_largeq.assign(n, 0);
    for (int i = 0; i < n; ++i) s.cptr[i + 1] = s.cptr[i] + (int)conns[i].size();
    s.cidx.resize(s.cptr[n]);
    for (int i = 0; i < n; ++i) {
        std::copy(conns[i].begin(), conns[i].end(), s.cidx.begin() + s.cptr[i]);
        auto it = std::find(conns[i].begin(), conns[i].end(), i);
        s.nconn_largeq[i] = (int)std::distance(it, conns[i].end());
    }
}

// Assembly::searchNonlocalNeighbors, include/assembly.h:117-130 (includes the particle itself)
static void search_nonlocal(Sys &s, double cutoff_ratio) {
    const int n = s.N;
    std::vector<std::vector<int>> nl(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (distanceTo(&s.xyz[3 * i], &s.xyz[3 * j]) < cutoff_ratio * (s.nonlocal_L[i])) nl[i].push_back(j);
    s.nlptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) s.nlptr[i + 1] = s.nlptr[i] + (int)nl[i].size();
    s.nlidx.resize(s.nlptr[n]);
    for (int i = 0; i < n; ++i) std::copy(nl[i].begin(), nl[i].end(), s.nlidx.begin() + s.nlptr[i]);
}

// ---------------------------------------------------------------- material setup
// Ce (C11, C12, C44) from (E, mu), e.g. particle_elastic.h:77-88
static void ce_from_E(bool plane_stress, double E, double mu, double Ce[3]) {
    if (!plane_stress) {
        Ce[0] = E * (1.0 - mu) / (1.0 + mu) / (1.0 - 2.0 * mu);
        Ce[1] = E * mu / (1.0 + mu) / (1.0 - 2.0 * mu);
        Ce[2] = E / 2.0 / (1.0 + mu);
    } else {
        Ce[0] = E / (1.0 - mu) / (1.0 + mu);
        Ce[1] = E * mu / (1.0 + mu) / (1.0 - mu);
        Ce[2] = E / 2.0 / (1.0 + mu);
    }
}

// KnTv = el_mapping * Ce (2D) or radius * el_mapping * Ce (3D), then per-bond Kn/Tv
// (particle_elastic.h:90-110). Eigen's column-major GEMV accumulates columns in order
// and applies the scalar factor to the finished sum.
static void set_kntv(Sys &s, int i, const double Ce[3]) {
    double KnTv[3];
    for (int r = 0; r < 3; ++r) {
        bool lit = s.elmap_literal != 0;
        double c = el_map(s.cell, lit, r, 0) * Ce[0];
        c = c + el_map(s.cell, lit, r, 1) * Ce[1];
        c = c + el_map(s.cell, lit, r, 2) * Ce[2];
        KnTv[r] = (s.cell.dim == 2) ? c : c * s.cell.radius;
    }
    for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b) {
        if (s.cell.lattice == Hexagon2D) { s.Kn[b] = KnTv[0]; s.Tv[b] = KnTv[1]; }
        else { s.Kn[b] = KnTv[s.layer[b]]; s.Tv[b] = KnTv[2]; }
    }
}

// mat[] layout per particle (10 doubles):
//  Elastic:        [0]=critical_bstrain
//  ElasticDamage:  [0]=k0 [1]=k1 [2]=damage_threshold [3]=comp_tensile_ratio
//  J2Plasticity:   [0]=E [1]=mu [2]=sigmay [3]=xi [4]=H [5]=A [6]=damage_threshold [7]=critical_bstrain
//  FatigueHCF:     [0]=A [1]=B [2]=d [3]=E [4]=mu [5]=damage_threshold [6]=fatigue_limit_ratio

// ---------------------------------------------------------------- per-step kernels
// Particle::updateBondsGeometry + Bond::updatebGeometry (particle.h:163-181, bond.h:30-39)
static void update_geometry(Sys &s) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.N; ++i) {
        for (int l = 0; l < NL; ++l) {
            int lo = l == 0 ? s.bptr[i] : s.bsplit[i], hi = l == 0 ? s.bsplit[i] : s.bptr[i + 1];
            double dt = 0, tdt = 0, sx = 0, sy = 0, sz = 0;
            for (int b = lo; b < hi; ++b) {
                const double *x1 = &s.xyz[3 * i], *x2 = &s.xyz[3 * s.p2[b]];
                s.dis[b] = distanceTo(x1, x2);
                bool broken = std::fabs(s.bdamage[b] - 1.0) < EPS;
                s.dL[b] = broken ? 0.0 : (s.dis[b] - s.dis0[b]);
                s.bstrain[b] = broken ? 0.0 : (s.dL[b] / s.dis[b]);
                s.dLe[b] = broken ? 0.0 : (s.dis[b] - s.dis0[b] - s.dLp[b]);
                s.cs[3 * b + 0] = broken ? 0.0 : ((x1[0] - x2[0]) / s.dis[b]);
                s.cs[3 * b + 1] = broken ? 0.0 : ((x1[1] - x2[1]) / s.dis[b]);
                s.cs[3 * b + 2] = broken ? 0.0 : ((x1[2] - x2[2]) / s.dis[b]);
                dt += s.dLe[b];
                tdt += s.Tv[b] * s.dLe[b];
                sx += s.cs[3 * b + 0]; sy += s.cs[3 * b + 1]; sz += s.cs[3 * b + 2];
            }
            s.dLe_total[2 * i + l] = dt; s.TdLe_total[2 * i + l] = tdt;
            s.cs_sum[6 * i + 3 * l + 0] = sx; s.cs_sum[6 * i + 3 * l + 1] = sy; s.cs_sum[6 * i + 3 * l + 2] = sz;
        }
    }
}

// updateBondsForce of one particle: particle_elastic.h:38-50 (Elastic);
// particle_elastic_damage.h:74-87, particle_j2plasticity.h:63-76, particle_fatigue_hcf.h:65-80
static void bonds_force_particle(Sys &s, int i) {
    for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b) {
        double f = 2. * s.Kn[b] * s.dLe[b] + 2. * s.Tv[b] * s.dLe_total[2 * i + s.layer[b]];
        if (s.ptype != Elastic)
            s.bdamage[b] = std::max(s.bdamage[b], std::max(s.damage[i], s.damage[s.p2[b]]));
        f *= (1.0 - s.bdamage[b]);
        s.bforce[b] = f;
    }
}

// Particle::updateParticleForce (particle.h:197-215)
static void particle_force(Sys &s, int i) {
    double P[3] = {0., 0., 0.};
    for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b) {
        int o = s.opp[b];
        P[0] += s.cs[3 * b + 0] * 0.5 * (s.bforce[b] + s.bforce[o]);
        P[1] += s.cs[3 * b + 1] * 0.5 * (s.bforce[b] + s.bforce[o]);
        P[2] += s.cs[3 * b + 2] * 0.5 * (s.bforce[b] + s.bforce[o]);
    }
    s.Pin[3 * i + 0] = P[0]; s.Pin[3 * i + 1] = P[1]; s.Pin[3 * i + 2] = P[2];
}

// Particle::updateParticleStress (particle.h:217-240)
static void particle_stress(Sys &s, int i) {
    double st[6] = {0, 0, 0, 0, 0, 0};
    double V_m = s.cell.volume * s.nb[i] / s.cell.nneighbors;
    for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b) {
        int o = s.opp[b];
        const double *c = &s.cs[3 * b];
        double ff = s.bforce[b] + s.bforce[o];
        st[0] += 0.5 / V_m * (s.dis[b]) * 0.5 * ff * (c[0]) * (c[0]);
        st[1] += 0.5 / V_m * (s.dis[b]) * 0.5 * ff * (c[1]) * (c[1]);
        st[2] += 0.5 / V_m * (s.dis[b]) * 0.5 * ff * (c[2]) * (c[2]);
        st[3] += 0.5 / V_m * (s.dis[b]) * 0.5 * ff * (c[1]) * (c[2]);
        st[4] += 0.5 / V_m * (s.dis[b]) * 0.5 * ff * (c[0]) * (c[2]);
        st[5] += 0.5 / V_m * (s.dis[b]) * 0.5 * ff * (c[0]) * (c[1]);
    }
    for (int k = 0; k < 6; ++k) s.stress[6 * i + k] = st[k];
}

// Particle::updateParticleDamageVisual (particle.h:150-161)
static void particle_damage_visual(Sys &s, int i) {
    double dv = 0;
    for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b) dv += s.bdamage[b];
    s.damage_visual[i] = dv / (double)s.nb[i];
}

// Assembly::updateForceState (assembly.h:178-191): all bond forces, then all particle sums
static void update_force_state(Sys &s) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.N; ++i) bonds_force_particle(s, i);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.N; ++i) {
        particle_force(s, i);
        particle_stress(s, i);
        particle_damage_visual(s, i);
    }
}

// Solver::updateRR (solver.h:108-128)
static void update_rr(Sys &s) {
    s.reaction.clear();
    for (int i = 0; i < s.N; ++i)
        for (int k = 0; k < s.dim; ++k) {
            s.residual[s.dim * i + k] = (1 - s.cons[3 * i + k]) * (s.Pex[3 * i + k] - s.Pin[3 * i + k]);
            if (s.cons[3 * i + k] == 1) s.reaction.push_back(s.Pin[3 * i + k]);
        }
}

static double norm2(const std::vector<double> &v) {
    double a = 0;
    for (double x : v) a += x * x;
    return std::sqrt(a);
}

// ---------------------------------------------------------------- stiffness
typedef std::array<std::array<double, NDIM>, NDIM> Blk;

// fdu2dxyz: diagonal block (derivative.h:7-45)
static Blk fdu2dxyz(const Sys &s, int i) {
    Blk d{};
    for (int l = 0; l < NL; ++l) {
        int lo = l == 0 ? s.bptr[i] : s.bsplit[i], hi = l == 0 ? s.bsplit[i] : s.bptr[i + 1];
        const double *S = &s.cs_sum[6 * i + 3 * l];
        for (int b = lo; b < hi; ++b) {
            const double *c = &s.cs[3 * b];
            double w = (1 - s.bdamage[b]), Kn = s.Kn[b], Tv = s.Tv[b];
            d[0][0] += w * ((Kn * c[0] + Tv * S[0]) * c[0] + (Kn + Tv) * std::pow(c[0], 2));
            d[0][1] += w * ((Kn * c[1] + Tv * S[1]) * c[0] + (Kn + Tv) * c[1] * c[0]);
            d[0][2] += w * ((Kn * c[2] + Tv * S[2]) * c[0] + (Kn + Tv) * c[2] * c[0]);
            d[1][1] += w * ((Kn * c[1] + Tv * S[1]) * c[1] + (Kn + Tv) * std::pow(c[1], 2));
            d[1][2] += w * ((Kn * c[2] + Tv * S[2]) * c[1] + (Kn + Tv) * c[2] * c[1]);
            d[2][2] += w * ((Kn * c[2] + Tv * S[2]) * c[2] + (Kn + Tv) * std::pow(c[2], 2));
        }
    }
    // derivative.h:32-39: [1][0] is a self-assignment in the reference (stays 0)
    if (!s.diag10_literal) d[1][0] = d[0][1];
    d[2][0] = d[0][2];
    d[2][1] = d[1][2];
    return d;
}

// product factor used by the direct term: pow(c_r,2) on the diagonal, c_max*c_min off it
static inline double cc_prod(double KT, const double *c, int r, int q) {
    if (r == q) return KT * std::pow(c[r], 2);
    int a = std::max(r, q), b = std::min(r, q);
    return KT * c[a] * c[b];
}

// fdu2dxyz1 / fdu2dxyz2: off-diagonal block for layer l (derivative.h:47-142 and 144-239)
static Blk fdu2dxyz_layer(const Sys &s, int i, int j, int l) {
    Blk d{};
    int lo = l == 0 ? s.bptr[i] : s.bsplit[i], hi = l == 0 ? s.bsplit[i] : s.bptr[i + 1];
    for (int b1 = lo; b1 < hi; ++b1) {
        if (s.p2[b1] != j) continue;
        for (int b2 = lo; b2 < hi; ++b2) {
            const double *c2 = &s.cs[3 * b2];
            if (s.p2[b1] == s.p2[b2]) {
                const double *S = &s.cs_sum[6 * s.p2[b2] + 3 * l];
                double w = (1 - s.bdamage[b2]), Kn = s.Kn[b2], Tv = s.Tv[b2];
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q)
                        d[r][q] += w * ((-Kn * c2[q] + Tv * S[q]) * c2[r] - cc_prod(Kn + Tv, c2, r, q));
            } else {
                const double *c1 = &s.cs[3 * b1];
                double a = 0.5 * ((1 - s.bdamage[b1]) * s.Tv[b1] + (1 - s.bdamage[b2]) * s.Tv[b2]);
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) d[r][q] -= a * c2[r] * c1[q];
                int k = s.p2[b2];
                int lo3 = l == 0 ? s.bptr[k] : s.bsplit[k], hi3 = l == 0 ? s.bsplit[k] : s.bptr[k + 1];
                for (int b3 = lo3; b3 < hi3; ++b3) {
                    if (s.p2[b3] == j) {
                        const double *c3 = &s.cs[3 * b3];
                        double a3 = 0.5 * ((1 - s.bdamage[b3]) * s.Tv[b3] + (1 - s.bdamage[b2]) * s.Tv[b2]);
                        for (int r = 0; r < 3; ++r)
                            for (int q = 0; q < 3; ++q) d[r][q] -= a3 * c2[r] * c3[q];
                    }
                }
            }
        }
        return d;
    }
    for (int b1 = lo; b1 < hi; ++b1) {
        int k = s.p2[b1];
        int lo2 = l == 0 ? s.bptr[k] : s.bsplit[k], hi2 = l == 0 ? s.bsplit[k] : s.bptr[k + 1];
        const double *c1 = &s.cs[3 * b1];
        for (int b2 = lo2; b2 < hi2; ++b2) {
            if (s.p2[b2] == j) {
                const double *c2 = &s.cs[3 * b2];
                double a = 0.5 * ((1 - s.bdamage[b2]) * s.Tv[b2] + (1 - s.bdamage[b1]) * s.Tv[b1]);
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) d[r][q] -= a * c1[r] * c2[q];
            }
        }
    }
    return d;
}

// Particle::hasAFEMneighbor (particle.h:242-256)
static bool has_afem_neighbor(const Sys &s, int i, int j, int l) {
    int lo = l == 0 ? s.bptr[i] : s.bsplit[i], hi = l == 0 ? s.bsplit[i] : s.bptr[i + 1];
    for (int b1 = lo; b1 < hi; ++b1) {
        if (s.p2[b1] == j) return true;
        int k = s.p2[b1];
        int lo2 = l == 0 ? s.bptr[k] : s.bsplit[k], hi2 = l == 0 ? s.bsplit[k] : s.bptr[k + 1];
        for (int b2 = lo2; b2 < hi2; ++b2)
            if (s.p2[b2] == j) return true;
    }
    return false;
}

// Stiffness::localStiffnessANA (stiffness.h:67-101)
static Blk local_stiffness_ana(const Sys &s, int i, int j) {
    bool f1 = false, f2 = false;
    Blk K{}, K1{}, K2{};
    if (has_afem_neighbor(s, i, j, 0)) { K1 = fdu2dxyz_layer(s, i, j, 0); f1 = true; }
    if (has_afem_neighbor(s, i, j, 1)) { K2 = fdu2dxyz_layer(s, i, j, 1); f2 = true; }
    if (i == j) K = fdu2dxyz(s, i);
    else if (f1 && f2) { for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) K[r][q] = K1[r][q] + K2[r][q]; }
    else if (f1) K = K1;
    else if (f2) K = K2;
    return K;
}

// ---- model hooks needed by the finite-difference block
static void update_state_var_particle(Sys &s, int i);

static void bonds_geometry_particle(Sys &s, int i) {
    for (int l = 0; l < NL; ++l) {
        int lo = l == 0 ? s.bptr[i] : s.bsplit[i], hi = l == 0 ? s.bsplit[i] : s.bptr[i + 1];
        double dt = 0, tdt = 0, sx = 0, sy = 0, sz = 0;
        for (int b = lo; b < hi; ++b) {
            const double *x1 = &s.xyz[3 * i], *x2 = &s.xyz[3 * s.p2[b]];
            s.dis[b] = distanceTo(x1, x2);
            bool broken = std::fabs(s.bdamage[b] - 1.0) < EPS;
            s.dL[b] = broken ? 0.0 : (s.dis[b] - s.dis0[b]);
            s.bstrain[b] = broken ? 0.0 : (s.dL[b] / s.dis[b]);
            s.dLe[b] = broken ? 0.0 : (s.dis[b] - s.dis0[b] - s.dLp[b]);
            s.cs[3 * b + 0] = broken ? 0.0 : ((x1[0] - x2[0]) / s.dis[b]);
            s.cs[3 * b + 1] = broken ? 0.0 : ((x1[1] - x2[1]) / s.dis[b]);
            s.cs[3 * b + 2] = broken ? 0.0 : ((x1[2] - x2[2]) / s.dis[b]);
            dt += s.dLe[b]; tdt += s.Tv[b] * s.dLe[b];
            sx += s.cs[3 * b + 0]; sy += s.cs[3 * b + 1]; sz += s.cs[3 * b + 2];
        }
        s.dLe_total[2 * i + l] = dt; s.TdLe_total[2 * i + l] = tdt;
        s.cs_sum[6 * i + 3 * l + 0] = sx; s.cs_sum[6 * i + 3 * l + 1] = sy; s.cs_sum[6 * i + 3 * l + 2] = sz;
    }
}

// Particle::resetParticleStateVariables (particle.h:128-148)
static void reset_particle_state(Sys &s, int i) {
    s.damage[i] = s.damage_last[i];
    for (int k = 0; k < s.nsv; ++k) s.sv[NSV * i + k] = s.sv_last[NSV * i + k];
    for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b) {
        s.dis[b] = s.dis_last[b]; s.bdamage[b] = s.bdamage_last[b];
        s.bforce[b] = s.bforce_last[b]; s.dLp[b] = s.dLp_last[b];
    }
}

// Stiffness::localStiffnessFD (stiffness.h:103-154) with resumeParticleState (particle.h:183-195).
// Literal: the restore computes (x + h) - h, which can leave x one ulp off, and later blocks see
// those positions. fd_exact = 1 restores x exactly instead (calc_stiffness then also starts from
// freshly evaluated forces), the order-independent quotient a parallel evaluation computes.
static Blk local_stiffness_fd(Sys &s, int i, int j) {
    std::vector<double> Kij;
    double xt[3] = {s.xyz[3 * j], s.xyz[3 * j + 1], s.xyz[3 * j + 2]};
    const double x_saved[3] = {xt[0], xt[1], xt[2]};
    particle_force(s, i);
    double Pt[3] = {s.Pin[3 * i], s.Pin[3 * i + 1], s.Pin[3 * i + 2]};
    const double h = EPS * s.cell.radius;
    for (int r = 0; r < s.dim; ++r) {
        xt[r] += EPS * s.cell.radius;
        for (int k = 0; k < 3; ++k) s.xyz[3 * j + k] = xt[k];
        for (int c = s.cptr[j]; c < s.cptr[j + 1]; ++c) bonds_geometry_particle(s, s.cidx[c]);
        for (int c = s.cptr[j]; c < s.cptr[j + 1]; ++c) update_state_var_particle(s, s.cidx[c]);
        for (int c = s.cptr[j]; c < s.cptr[j + 1]; ++c) bonds_force_particle(s, s.cidx[c]);
        particle_force(s, i);
        for (int q = 0; q < s.dim; ++q) Kij.push_back((s.Pin[3 * i + q] - Pt[q]) / EPS / (s.cell.radius));
        if (s.fd_exact) xt[r] = x_saved[r];
        else xt[r] -= EPS * s.cell.radius;
        for (int k = 0; k < 3; ++k) s.xyz[3 * j + k] = xt[k];
        for (int c = s.cptr[j]; c < s.cptr[j + 1]; ++c) {
            bonds_geometry_particle(s, s.cidx[c]);
            reset_particle_state(s, s.cidx[c]);
        }
        for (int c = s.cptr[j]; c < s.cptr[j + 1]; ++c) bonds_force_particle(s, s.cidx[c]);
    }
    (void)h;
    Blk K{};
    for (int r = 0; r < s.dim; ++r)
        for (int q = 0; q < s.dim; ++q) K[r][q] = Kij[s.dim * q + r];
    return K;
}

// Stiffness::calcStiffness (stiffness.h:228-263): block (i, j in conns(i)) at rows i*dim+r, cols j*dim+s
static void calc_stiffness(Sys &s) {
    const int d = s.dim, n = s.N;
    s.krow.assign(d * n + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int r = 0; r < d; ++r) s.krow[d * i + r + 1] = (s.cptr[i + 1] - s.cptr[i]) * d;
    for (int R = 0; R < d * n; ++R) s.krow[R + 1] += s.krow[R];
    s.kcol.assign(s.krow[d * n], 0);
    s.kval.assign(s.krow[d * n], 0.0);
    auto fill = [&](int i) {
        int nc = s.cptr[i + 1] - s.cptr[i];
        for (int c = 0; c < nc; ++c) {
            int j = s.cidx[s.cptr[i] + c];
            Blk K = s.stiffness_fd ? local_stiffness_fd(s, i, j) : local_stiffness_ana(s, i, j);
            for (int r = 0; r < d; ++r)
                for (int q = 0; q < d; ++q) {
                    int pos = s.krow[d * i + r] + c * d + q;
                    s.kcol[pos] = j * d + q;
                    s.kval[pos] = K[r][q];
                }
        }
    };
    if (s.stiffness_fd) {
        if (s.fd_exact) {  // base forces at the current positions
            update_geometry(s);
            for (int i = 0; i < n; ++i) bonds_force_particle(s, i);
        }
        for (int i = 0; i < n; ++i) fill(i);  // FD mutates neighbour state: serial, as the reference
    } else {
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) fill(i);
    }
}

// Stiffness::updateStiffnessDispBC (stiffness.h:265-294). Explicit zeros that the
// reference inserts outside the pattern do not change any product, so they are not stored.
static void update_stiffness_disp_bc(Sys &s) {
    const int d = s.dim, n = s.N, nd = d * n;
    double nrm = 0;
    for (int R = 0; R < nd; ++R) {
        double v = 0;
        for (int p = s.krow[R]; p < s.krow[R + 1]; ++p) if (s.kcol[p] == R) v = s.kval[p];
        nrm += v * v;
    }
    nrm = std::sqrt(nrm);
    std::vector<char> elim(nd, 0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < d; ++k)
            if (s.cons[3 * i + k] == 1 || std::fabs(s.damage_visual[i] - 1.0) < EPS) elim[d * i + k] = 1;
    for (int R = 0; R < nd; ++R)
        for (int p = s.krow[R]; p < s.krow[R + 1]; ++p) {
            int C = s.kcol[p];
            if (elim[R] || elim[C]) s.kval[p] = 0.0;
            if (elim[R] && C == R) s.kval[p] = nrm;
        }
}

// Eigen::ConjugateGradient<SparseMatrix<double>, Lower|Upper> + DiagonalPreconditioner
// (restated from Eigen 3.4: ConjugateGradient.h conjugate_gradient(), BasicPreconditioners.h).
// With Lower|Upper and a column-major matrix Eigen multiplies by the transposed
// (row-major) view, so A = K^T when cg_transpose is set (solver.h:195).
static int cg_solve(Sys &s) {
    const int nd = s.dim * s.N;
    // A in CSR: K as stored, or its transpose built here
    std::vector<int> art, act; std::vector<double> avt;
    if (s.cg_transpose) {
        art.assign(nd + 1, 0);
        for (size_t p = 0; p < s.kcol.size(); ++p) art[s.kcol[p] + 1]++;
        for (int R = 0; R < nd; ++R) art[R + 1] += art[R];
        act.resize(s.kcol.size()); avt.resize(s.kcol.size());
        std::vector<int> fill(art.begin(), art.end() - 1);
        for (int R = 0; R < nd; ++R)
            for (int p = s.krow[R]; p < s.krow[R + 1]; ++p) { int q = fill[s.kcol[p]]++; act[q] = R; avt[q] = s.kval[p]; }
    }
    const std::vector<int> &ar = s.cg_transpose ? art : s.krow, &ac = s.cg_transpose ? act : s.kcol;
    const std::vector<double> &av = s.cg_transpose ? avt : s.kval;
    std::vector<double> invdiag(nd, 1.0);
    for (int R = 0; R < nd; ++R)
        for (int p = ar[R]; p < ar[R + 1]; ++p)
            if (ac[p] == R) { if (av[p] != 0.0) invdiag[R] = 1.0 / av[p]; break; }
    auto spmv = [&](const std::vector<double> &x, std::vector<double> &y) {
#pragma omp parallel for schedule(static)
        for (int R = 0; R < nd; ++R) {
            double t = 0;
            for (int p = ar[R]; p < ar[R + 1]; ++p) t += av[p] * x[ac[p]];
            y[R] = t;
        }
    };
    auto dot = [&](const std::vector<double> &a, const std::vector<double> &b) {
        double t = 0;
        for (int R = 0; R < nd; ++R) t += a[R] * b[R];
        return t;
    };
    const std::vector<double> &rhs = s.residual;
    std::vector<double> &x = s.disp;
    x.assign(nd, 0.0);
    long maxIters = s.cg_maxit < 0 ? 2L * nd : s.cg_maxit;
    double tol = s.cg_tol;
    std::vector<double> r(rhs), p(nd), z(nd), t(nd);
    double rhsNorm2 = dot(rhs, rhs);
    if (rhsNorm2 == 0) { s.cg_iters = 0; s.cg_error = 0; s.cg_success = 1; return 1; }
    const double considerAsZero = 2.2250738585072014e-308;  // DBL_MIN
    double threshold = std::max(tol * tol * rhsNorm2, considerAsZero);
    double residualNorm2 = dot(r, r);
    if (residualNorm2 < threshold) {
        s.cg_iters = 0; s.cg_error = std::sqrt(residualNorm2 / rhsNorm2); s.cg_success = s.cg_error <= tol; return s.cg_success;
    }
    for (int R = 0; R < nd; ++R) p[R] = invdiag[R] * r[R];
    double absNew = dot(r, p);
    long i = 0;
    while (i < maxIters) {
        spmv(p, t);
        double alpha = absNew / dot(p, t);
        for (int R = 0; R < nd; ++R) x[R] += alpha * p[R];
        for (int R = 0; R < nd; ++R) r[R] -= alpha * t[R];
        residualNorm2 = dot(r, r);
        if (residualNorm2 < threshold) break;
        for (int R = 0; R < nd; ++R) z[R] = invdiag[R] * r[R];
        double absOld = absNew;
        absNew = dot(r, z);
        double beta = absNew / absOld;
        for (int R = 0; R < nd; ++R) p[R] = z[R] + beta * p[R];
        i++;
    }
    s.cg_error = std::sqrt(residualNorm2 / rhsNorm2);
    s.cg_iters = (int)i;
    s.cg_success = s.cg_error <= tol;
    s.cg_log.push_back(s.cg_iters);
    return s.cg_success;
}

// Solver::solveLinearSystem (solver.h:91-106): every particle moves by its DoFs of disp
static void solve_linear_system(Sys &s) {
    cg_solve(s);
    for (int i = 0; i < s.N; ++i)
        for (int k = 0; k < s.dim; ++k) s.xyz[3 * i + k] = s.xyz[3 * i + k] + s.disp[s.dim * i + k];
}

// Solver::NewtonIteration (solver.h:53-89)
static int newton_iteration(Sys &s) {
    update_geometry(s);
    update_force_state(s);
    update_rr(s);
    double norm_residual = norm2(s.residual);
    double norm_reaction = norm2(s.reaction);
    double tol_multiplier = std::max(norm_residual, norm_reaction);
    if (s.verbose) printf("|  Norm of residual is %.5e, norm of reaction is %.5e\n", norm_residual, norm_reaction);
    int ni = 0;
    while (norm_residual > s.tol_NR * tol_multiplier) {
        if (++ni > s.max_NR_iter) { s.last_newton = s.max_NR_iter; return s.max_NR_iter; }
        const double cg_tol_set = s.cg_tol;
        if (s.cg_forcing)  // the reduction still needed to pass the Newton test (factor 2 margin), in [cg_tol, 0.1]
            s.cg_tol = std::max(cg_tol_set, std::min(0.1, 0.5 * s.tol_NR * tol_multiplier / norm_residual));
        solve_linear_system(s);
        s.cg_tol = cg_tol_set;
        update_geometry(s);
        update_force_state(s);
        update_rr(s);
        norm_residual = norm2(s.residual);
        if (s.verbose) printf("|  |  Iteration-%d: residual %.3e ratio %.3e (cg %d)\n", ni, norm_residual, norm_residual / tol_multiplier, s.cg_iters);
    }
    s.last_newton = ni;
    return ni;
}

// ---------------------------------------------------------------- state updates
// ParticleElasticDamage::calcStrainEnergyDensity (particle_elastic_damage.h:89-104);
// the same sum is ParticleFatigueHCF::calcEqEnergy's returned energy_total (particle_fatigue_hcf.h:82-125)
static double strain_energy(const Sys &s, int i) {
    double V_m = s.cell.volume * s.nb[i] / s.cell.nneighbors;
    double e = 0;
    for (int l = 0; l < NL; ++l) {
        int lo = l == 0 ? s.bptr[i] : s.bsplit[i], hi = l == 0 ? s.bsplit[i] : s.bptr[i + 1];
        for (int b = lo; b < hi; ++b) e += 0.5 * (s.Kn[b] * s.dLe[b] * s.dLe[b]) / V_m;
        e += 0.5 * s.TdLe_total[2 * i + l] * s.dLe_total[2 * i + l] / V_m;
    }
    return e;
}

// updateParticleStateVariables for one particle
static void update_state_var_particle(Sys &s, int i) {
    double *sv = &s.sv[NSV * i];
    const double *sl = &s.sv_last[NSV * i];
    const double *m = &s.mat[10 * i];
    if (s.ptype == ElasticDamage) {  // particle_elastic_damage.h:150-161
        double k0 = m[0], k1 = m[1];
        sv[0] = strain_energy(s, i);
        double delta = sv[0] - sl[0];
        double f = sv[0] - k1 * s.damage[i] - k0;
        if (f > 0 && delta > 0) s.Ddot_local[i] = (1 / k1) * delta;
        else s.Ddot_local[i] = 0;
    } else if (s.ptype == J2Plasticity) {  // particle_j2plasticity.h:78-140
        double E = m[0], mu = m[1], sigmay = m[2], xi = m[3], H = m[4], A = m[5];
        double sigma_m = 0.0, sigma_eq = 0.0, triax = 0.0, dlambda = 0.0;
        double st[6], dpl[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 6; ++k) st[k] = s.stress[6 * i + k];
        sigma_m = 1.0 / 3.0 * (st[0] + st[1] + st[2]);
        for (int k = 0; k < 3; ++k) st[k] -= sigma_m;
        for (int k = 0; k < 6; ++k) st[k] -= sv[k + 1];
        for (int k = 0; k < 6; ++k) {
            if (k < 3) sigma_eq += st[k] * st[k];
            else sigma_eq += 2.0 * st[k] * st[k];
        }
        sigma_eq = std::sqrt(3.0 / 2.0 * sigma_eq);
        double yield_func = sigma_eq - (sigmay + (1.0 - xi) * H * sv[0]);
        double G = E / 2 / (1 + mu);
        if (yield_func > 0.0) dlambda = yield_func / (3 * G + H);
        sv[0] += dlambda;
        for (int k = 0; k < 6; ++k) {
            if (std::fabs(sigma_eq) > EPS) {
                dpl[k] = dlambda * 1.5 * st[k] / sigma_eq;
                sv[k + 1] += 2. / 3. * xi * H * dpl[k];
            }
        }
        for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b) {
            const double *c = &s.cs[3 * b];
            double ddLp = s.dis[b] * (dpl[0] * c[0] * c[0] + dpl[1] * c[1] * c[1] + dpl[2] * c[2] * c[2] +
                                      2 * dpl[3] * c[1] * c[2] + 2 * dpl[4] * c[0] * c[2] + 2 * dpl[5] * c[0] * c[1]);
            s.dLp[b] += ddLp;
        }
        if (sigma_eq > EPS) triax = sigma_m / sigma_eq;
        s.Ddot_local[i] = dlambda * (1.0 + A * triax);
    } else if (s.ptype == FatigueHCF) {  // particle_fatigue_hcf.h:127-145
        double A = m[0], B = m[1], dd = m[2];
        double curr = strain_energy(s, i);
        sv[1] = std::min(curr, sv[1]);
        sv[2] = std::max(curr, sv[2]);
        sv[0] = std::max(sv[2], sv[0]);
        if (sv[2] - sv[1] > 0) s.Ddot_local[i] = A * std::pow(sv[0], B) * std::pow(sv[2] - sv[1], dd);
        else s.Ddot_local[i] = 0;
    }
}

// Assembly::updateStateVar (assembly.h:161-166)
static void update_state_var(Sys &s) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.N; ++i) update_state_var_particle(s, i);
}

// nonlocal Ddot average: SolverStatic::updateStaticDamage (solver_static.h:31-52) and
// SolverFatigue::updateFatigueDamage (solver_fatigue.h:75-114)
static void nonlocal_ddot(Sys &s, int undamaged) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.N; ++i) {
        if (s.type[i] == undamaged) continue;
        double A = 0, dn = 0;
        for (int q = s.nlptr.empty() ? 0 : s.nlptr[i]; q < (s.nlptr.empty() ? 0 : s.nlptr[i + 1]); ++q) {
            int j = s.nlidx[q];
            if (s.type[j] != undamaged) {
                double dis = distanceTo(&s.xyz[3 * i], &s.xyz[3 * j]);
                double V_m = s.cell.volume * s.nb[j] / s.cell.nneighbors;
                dn += s.Ddot_local[j] * func_phi(dis, s.nonlocal_L[i]) * V_m;
                A += func_phi(dis, s.nonlocal_L[i]) * V_m;
            }
        }
        dn /= A;
        s.Ddot_nonlocal[i] = dn;
    }
}

// updateParticleStaticDamage for one particle
static bool static_damage_particle(Sys &s, int i) {
    const double *m = &s.mat[10 * i];
    if (s.ptype == Elastic) {  // particle_elastic.h:52-68
        bool any = false;
        double crit = m[0];
        for (int b = s.bptr[i]; b < s.bptr[i + 1]; ++b)
            if (s.bstrain[b] >= crit && std::fabs(s.bdamage[b] - 1.0) > EPS) { any = true; s.bdamage[b] = 1; }
        return any;
    }
    double thr = (s.ptype == ElasticDamage) ? m[2] : (s.ptype == J2Plasticity ? m[6] : 0.0);
    if (s.ptype == ElasticDamage || s.ptype == J2Plasticity) {  // particle_elastic_damage.h:163-173, particle_j2plasticity.h:142-152
        s.damage[i] += s.Ddot_nonlocal[i];
        if (s.damage[i] >= thr) s.damage[i] = 1;
        return std::fabs(s.damage[i] - s.damage_last[i]) > 1e-3;
    }
    return false;  // FatigueHCF has no static damage (particle.h:63 default)
}

// SolverStatic::updateStaticDamage (solver_static.h:31-62)
static bool update_static_damage(Sys &s, int undamaged) {
    nonlocal_ddot(s, undamaged);
    bool any = false;
    for (int i = 0; i < s.N; ++i)
        if (s.type[i] != undamaged) any = static_damage_particle(s, i) || any;
    return any;
}

// ParticleFatigueHCF::updateParticleFatigueDamage (particle_fatigue_hcf.h:147-160)
static bool fatigue_damage_particle(Sys &s, int i, double dt) {
    if (s.ptype != FatigueHCF) return false;  // particle.h:64 default
    const double *m = &s.mat[10 * i];
    double *sv = &s.sv[NSV * i];
    s.damage[i] += s.Ddot_nonlocal[i] * dt;
    if (s.damage[i] >= m[5]) s.damage[i] = 1;
    sv[1] = INT_MAX;
    sv[2] = 0;
    return std::fabs(s.damage[i] - s.damage_last[i]) > 1e-3;
}

// SolverFatigue::updateFatigueDamage (solver_fatigue.h:75-126)
static bool update_fatigue_damage(Sys &s, double dNdt, int undamaged) {
    nonlocal_ddot(s, undamaged);
    bool any = false;
    for (int i = 0; i < s.N; ++i)
        if (s.type[i] != undamaged) any = fatigue_damage_particle(s, i, dNdt) || any;
    return any;
}

// Assembly::storeStateVar / resetStateVar (assembly.h:140-159; particle.h:108-148)
static void store_state(Sys &s) {
    s.xyz_last = s.xyz;
    s.damage_last = s.damage;
    s.sv_last = s.sv;
    s.dis_last = s.dis; s.bdamage_last = s.bdamage; s.bforce_last = s.bforce; s.dLp_last = s.dLp;
}
static void reset_state(Sys &s, bool reset_xyz) {
    if (reset_xyz) s.xyz = s.xyz_last;
    s.damage = s.damage_last;
    s.sv = s.sv_last;
    s.dis = s.dis_last; s.bdamage = s.bdamage_last; s.bforce = s.bforce_last; s.dLp = s.dLp_last;
}

// ---------------------------------------------------------------- boundary conditions
// Solver::updateDisplacementBC (solver.h:130-166)
static void apply_disp_bc(Sys &s, const LoadStep &ls) {
    for (const auto &bc : ls.disp) {
        int ax = bc.flag == 'x' ? 0 : (bc.flag == 'y' ? 1 : (bc.flag == 'z' ? 2 : -1));
        if (ax < 0) continue;
        for (int i : bc.group) {
            double d[3] = {0, 0, 0};
            d[ax] = bc.step;
            if (bc.mode == Relative) for (int k = 0; k < 3; ++k) s.xyz[3 * i + k] = s.xyz[3 * i + k] + d[k];
            else for (int k = 0; k < 3; ++k) s.xyz[3 * i + k] = d[k];
            s.cons[3 * i + ax] = 1;
        }
    }
}
// Solver::updateForceBC (solver.h:168-190)
static void apply_force_bc(Sys &s, const LoadStep &ls) {
    for (const auto &bc : ls.force) {
        int num = (int)bc.group.size();
        for (int i : bc.group) {
            if (bc.mode == Relative) {
                s.Pex[3 * i + 0] += bc.fx / num; s.Pex[3 * i + 1] += bc.fy / num; s.Pex[3 * i + 2] += bc.fz / num;
            } else {
                s.Pex[3 * i + 0] = bc.fx / num; s.Pex[3 * i + 1] = bc.fy / num; s.Pex[3 * i + 2] = bc.fz / num;
            }
        }
    }
}

// ---------------------------------------------------------------- solvers
// SolverStatic::solveProblemStep (solver_static.h:97-149)
static bool solve_step_static(Sys &s, const LoadStep &ls, int undamaged) {
    apply_force_bc(s, ls);
    apply_disp_bc(s, ls);
    int n_newton = 0;
    bool new_damaged = false;
    do {
        calc_stiffness(s);
        if (s.dirichlet_static) update_stiffness_disp_bc(s);  // commented out in the reference (:113)
        n_newton = newton_iteration(s);
        s.newton_log.push_back(n_newton);
        if (n_newton >= s.max_NR_iter) break;
        update_state_var(s);
        new_damaged = update_static_damage(s, undamaged);
        store_state(s);
        if (new_damaged) { update_geometry(s); update_force_state(s); }
    } while (new_damaged);
    return n_newton < s.max_NR_iter;
}

// SolverStatic::solveProblem (solver_static.h:64-95) without the dump writer
static void solve_problem_static(Sys &s, std::vector<LoadStep> load, int undamaged) {
    int n_step = (int)load.size();
    for (int i = 0; i < n_step; i++) {
        for (;;) {
            std::vector<double> pex_saved = s.Pex;
            bool ok = solve_step_static(s, load[i], undamaged);
            if (ok) break;
            reset_state(s, true);
            if (s.rollback_pex) s.Pex = pex_saved;
            load[i].cutHalf();
            load.insert(load.begin() + i, load[i]);
            ++n_step;
            ++s.n_restarts;
        }
    }
    s.n_steps_done = n_step;
}

// SolverFatigue::solveProblemStep (solver_fatigue.h:277-298)
static bool solve_step_fatigue(Sys &s, const LoadStep &ls) {
    apply_force_bc(s, ls);
    apply_disp_bc(s, ls);
    calc_stiffness(s);
    update_stiffness_disp_bc(s);
    int n_newton = newton_iteration(s);
    s.newton_log.push_back(n_newton);
    return n_newton < s.max_NR_iter;
}

// SolverFatigue::solveProblemStatic (solver_fatigue.h:243-275) without the dump writer
static void solve_problem_fatigue_static(Sys &s, std::vector<LoadStep> load) {
    int n_step = (int)load.size();
    for (int i = 0; i < n_step; i++) {
        for (;;) {
            std::vector<double> pex_saved = s.Pex;
            bool ok = solve_step_fatigue(s, load[i]);
            if (ok) break;
            reset_state(s, true);
            if (s.rollback_pex) s.Pex = pex_saved;
            load[i].cutHalf();
            load.insert(load.begin() + i, load[i]);
            ++n_step;
            ++s.n_restarts;
        }
        update_state_var(s);
        store_state(s);
    }
    s.n_steps_done = n_step;
}

// SolverFatigue::solveProblemOneCycle (solver_fatigue.h:220-241)
static void solve_one_cycle(Sys &s, std::vector<LoadStep> &load, double dNdt, int undamaged) {
    for (auto &ls : load) {
        solve_step_fatigue(s, ls);
        update_state_var(s);
    }
    update_fatigue_damage(s, dNdt, undamaged);
    store_state(s);
    update_geometry(s);
    update_force_state(s);
}

// load templates, include/load_step_fatigue.h:22-85
static LoadStep make_fatigue_step(int ltype, double v, const std::vector<std::vector<int>> &g) {
    LoadStep ls;
    if (ltype == 0) {  // LoadUniaxialDisp
        ls.disp.push_back({g[0], Relative, 'x', 0.0});
        ls.disp.push_back({g[1], Relative, 'y', v});
        ls.disp.push_back({g[2], Relative, 'y', 0.0});
    } else if (ltype == 1) {  // LoadUniaxialForce
        ls.disp.push_back({g[0], Relative, 'x', 0.0});
        ls.force.push_back({g[1], Relative, 0.0, v, 0.0});
        ls.disp.push_back({g[2], Relative, 'y', 0.0});
    } else if (ltype == 2) {  // LoadDogBoneDisp(0, v, 0)
        ls.disp.push_back({g[0], Relative, 'x', 0});
        ls.disp.push_back({g[0], Relative, 'y', v});
        ls.disp.push_back({g[0], Relative, 'z', 0});
        ls.disp.push_back({g[1], Relative, 'x', 0.0});
        ls.disp.push_back({g[1], Relative, 'y', 0.0});
        ls.disp.push_back({g[1], Relative, 'z', 0.0});
    } else {  // LoadCTForce
        ls.disp.push_back({g[0], Relative, 'y', 0});
        ls.disp.push_back({g[1], Relative, 'x', 0.0});
        ls.disp.push_back({g[2], Relative, 'x', 0.0});
        ls.force.push_back({g[1], Relative, 0.0, v, 0.0});
        ls.force.push_back({g[2], Relative, 0.0, -v, 0.0});
    }
    return ls;
}

// SolverFatigue::solveProblemCyclic + generateLoadCycle (solver_fatigue.h:128-218)
static int solve_problem_cyclic(Sys &s, int ltype, int n_interval, const std::vector<double> &spec,
                                const std::vector<std::vector<int>> &groups, int tmode, double tau, int undamaged,
                                int max_cycles_run) {
    double n_cycles = 0.5 * (spec.size() - 1);
    double t = 0, dNdt = 1;
    int N = 0, ran = 0;
    do {
        if ((size_t)(2 * N + 2) >= spec.size()) return -1;  // out of range in the reference (UB)
        double f0 = spec[2 * N], f1 = spec[2 * N + 1], f2 = spec[2 * N + 2];
        std::vector<LoadStep> cyc;
        for (int k = 0; k < n_interval; ++k) cyc.push_back(make_fatigue_step(ltype, (f1 - f0) / n_interval, groups));
        for (int k = 0; k < n_interval; ++k) cyc.push_back(make_fatigue_step(ltype, (f2 - f1) / n_interval, groups));
        dNdt = (tmode == 0) ? (int)(1 / tau) : (int)(1 / tau * std::exp(t++ / tau));
        solve_one_cycle(s, cyc, dNdt, undamaged);
        N += (int)dNdt;
        ++ran;
        if (max_cycles_run > 0 && ran >= max_cycles_run) break;
    } while (N < n_cycles);
    return N;
}

// Assembly::writeDump (assembly.h:309-335)
static void write_dump(const Sys &s, const char *path, int step) {
    FILE *f = std::fopen(path, "a+");
    if (!f) return;
    std::fprintf(f, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n", step, s.N);
    std::fprintf(f, "%8.8f %8.8f\n", s.box[0], s.box[1]);
    std::fprintf(f, "%8.8f %8.8f\n", s.box[2], s.box[3]);
    std::fprintf(f, "%8.8f %8.8f\n", s.box[4], s.box[5]);
    std::fprintf(f, "ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 damage damage_visual state_var[0] \n");
    for (int i = 0; i < s.N; ++i) {
        const double *x = &s.xyz[3 * i], *x0 = &s.xyz0[3 * i], *st = &s.stress[6 * i];
        std::fprintf(f, "%d %d %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e\n", i, s.type[i],
                     x[0], x[1], x[2], x[0] - x0[0], x[1] - x0[1], x[2] - x0[2], st[0], st[1], st[2], st[3], st[4], st[5],
                     s.damage[i], s.damage_visual[i], s.sv[NSV * i]);
    }
    std::fclose(f);
}

// ---------------------------------------------------------------- lattice generators
// createRMatrix (utilities.h:8-82); row-major R[3][3]
static void rmatrix(int flag, const double ang[3], double R[9]) {
    double a1 = ang[0], a2 = ang[1], a3 = ang[2];
    for (int k = 0; k < 9; ++k) R[k] = 0;
    if (flag == 0) {
        R[0] = std::cos(a2) * std::cos(a3);
        R[1] = -std::cos(a2) * std::sin(a3);
        R[2] = std::sin(a2);
        R[3] = std::cos(a1) * std::sin(a3) + std::sin(a1) * std::sin(a2) * std::cos(a3);
        R[4] = std::cos(a1) * std::cos(a3) - std::sin(a1) * std::sin(a2) * std::sin(a3);
        R[5] = -std::sin(a1) * std::cos(a2);
        R[6] = std::sin(a1) * std::sin(a3) - std::cos(a1) * std::sin(a2) * std::cos(a3);
        R[7] = std::sin(a1) * std::cos(a3) + std::cos(a1) * std::sin(a2) * std::sin(a3);
        R[8] = std::cos(a1) * std::cos(a2);
    } else if (flag == 1) {
        R[0] = -std::cos(a1) * std::cos(a2) * std::cos(a3) - std::sin(a1) * std::sin(a3);
        R[1] = std::cos(a3) * std::sin(a1) - std::sin(a3) * std::cos(a2) * std::cos(a1);
        R[2] = std::sin(a2) * std::cos(a1);
        R[3] = std::sin(a3) * std::cos(a1) - std::cos(a3) * std::cos(a2) * std::sin(a1);
        R[4] = -std::cos(a1) * std::cos(a3) - std::sin(a1) * std::cos(a2) * std::sin(a3);
        R[5] = std::sin(a2) * std::sin(a1);
        R[6] = std::cos(a3) * std::sin(a2);
        R[7] = std::sin(a3) * std::sin(a2);
        R[8] = std::cos(a2);
    } else if (flag == 2) {
        R[0] = -std::sin(a1) * std::cos(a2) * std::sin(a3) + std::cos(a1) * std::cos(a3);
        R[1] = -std::sin(a3) * std::cos(a1) - std::cos(a3) * std::cos(a2) * std::sin(a1);
        R[2] = std::sin(a1) * std::sin(a2);
        R[3] = std::cos(a3) * std::sin(a1) + std::sin(a3) * std::cos(a2) * std::cos(a1);
        R[4] = -std::sin(a1) * std::sin(a3) + std::cos(a1) * std::cos(a2) * std::cos(a3);
        R[5] = -std::cos(a1) * std::sin(a2);
        R[6] = std::sin(a2) * std::sin(a3);
        R[7] = std::sin(a2) * std::cos(a3);
        R[8] = std::cos(a2);
    }
}

// Eigen GEMV R * p (column-major accumulation, columns in order)
static void rmul(const double R[9], const double p[3], double o[3]) {
    for (int r = 0; r < 3; ++r) {
        double c = R[3 * r + 0] * p[0];
        c = c + R[3 * r + 1] * p[1];
        c = c + R[3 * r + 2] * p[2];
        o[r] = c;
    }
}

static bool in_box2(const double *p, const double *box) {
    return p[0] >= box[0] && p[0] <= box[1] && p[1] >= box[2] && p[1] <= box[3];
}
static bool in_box3(const double *p, const double *box) {
    return in_box2(p, box) && p[2] >= box[4] && p[2] <= box[5];
}

// createPlateSQ2D (utilities.h:84-116), createPlateHEX2D (118-168), createCuboidSC3D (170-211),
// createCuboidFCC3D (213-301)
static std::vector<double> generate(int lattice, const double *box, double radius, const double R[9]) {
    std::vector<double> out;
    double p[3] = {0, 0, 0}, q[3];
    if (lattice == Square2D) {
        double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2));
        double bt[6] = {-a, a, -a, a, -a, a};
        double hx = 2 * radius, hy = hx;
        int pfr = (int)std::floor((bt[1] - bt[0]) / hx);
        int rows = (int)std::floor((bt[3] - bt[2]) / hy);
        for (int j = 1; j <= rows; j++) {
            p[1] = bt[2] + hy * (j - 1) + radius;
            for (int i = 1; i <= pfr; i++) {
                p[0] = bt[0] + hx * (i - 1) + radius;
                rmul(R, p, q);
                if (in_box2(q, box)) { out.push_back(q[0]); out.push_back(q[1]); out.push_back(0); }
            }
        }
    } else if (lattice == Hexagon2D) {
        double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2));
        double bt[6] = {-a, a, -a, a, -a, a};
        double hx = 2 * radius, hy = hx * std::sqrt(3.0) / 2.0;
        int pfr = 1 + (int)std::floor((bt[1] - bt[0]) / hx);
        int rows = 1 + (int)std::floor((bt[3] - bt[2]) / hy);
        std::vector<double> raw;
        for (int j = 1; j <= rows; j++) {
            p[1] = bt[2] + hy * (j - 1);
            if (j % 2 == 1)
                for (int i = 1; i <= pfr; i++) { p[0] = bt[0] + hx * (i - 1); raw.push_back(p[0]); raw.push_back(p[1]); raw.push_back(0); }
            else
                for (int i = 1; i <= pfr - 1; i++) { p[0] = bt[0] + hx * (2 * i - 1) / 2.0; raw.push_back(p[0]); raw.push_back(p[1]); raw.push_back(0); }
        }
        for (size_t k = 0; k < raw.size(); k += 3) {
            rmul(R, &raw[k], q);
            if (in_box2(q, box)) { out.push_back(q[0]); out.push_back(q[1]); out.push_back(q[2]); }
        }
    } else if (lattice == SimpleCubic3D) {
        double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2) + std::pow(box[5] - box[4], 2));
        double bt[6] = {-a, a, -a, a, -a, a};
        double hx = 2 * radius, hy = hx, hz = hx;
        int pfr = 1 + (int)std::floor((bt[1] - bt[0]) / hx);
        int rows = 1 + (int)std::floor((bt[3] - bt[2]) / hy);
        int layers = 1 + (int)std::floor((bt[5] - bt[4]) / hz);
        for (int k = 1; k <= layers; k++) {
            p[2] = bt[4] + hz * (k - 1);
            for (int j = 1; j <= rows; j++) {
                p[1] = bt[2] + hy * (j - 1);
                for (int i = 1; i <= pfr; i++) {
                    p[0] = box[0] + hx * (i - 1);  // uses box[0], not box_t[0] (utilities.h:195)
                    rmul(R, p, q);
                    if (in_box3(q, box)) { out.push_back(q[0]); out.push_back(q[1]); out.push_back(q[2]); }
                }
            }
        }
    } else if (lattice == FCC3D) {
        double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2) + std::pow(box[5] - box[4], 2));
        double bt[6] = {-a, a, -a, a, -a, a};
        double hx = 4 * radius / std::sqrt(2.0), hy = hx, hz = hx;
        int pfr = 1 + (int)std::floor((bt[1] - bt[0]) / hx);
        int rows = 1 + (int)std::floor((bt[3] - bt[2]) / (hy / 2.0));
        int layers = 1 + (int)std::floor((bt[5] - bt[4]) / (hz / 2.0));
        std::vector<double> raw;
        for (int k = 1; k <= layers; k++) {
            p[2] = bt[4] + hz / 2.0 * (k - 1);
            for (int j = 1; j <= rows; j++) {
                p[1] = bt[2] + hy / 2.0 * (j - 1);
                bool full = ((k % 2 == 1) == (j % 2 == 1));
                if (full)
                    for (int i = 1; i <= pfr; i++) { p[0] = bt[0] + hx * (i - 1); raw.push_back(p[0]); raw.push_back(p[1]); raw.push_back(p[2]); }
                else
                    for (int i = 1; i <= pfr - 1; i++) { p[0] = bt[0] + hx * (i - 1) + hx / 2.0; raw.push_back(p[0]); raw.push_back(p[1]); raw.push_back(p[2]); }
            }
        }
        for (size_t k = 0; k < raw.size(); k += 3) {
            rmul(R, &raw[k], q);
            if (in_box3(q, box)) { out.push_back(q[0]); out.push_back(q[1]); out.push_back(q[2]); }
        }
    }
    return out;
}

}  // namespace ora

// =================================================================== C API
using namespace ora;

extern "C" {

int ora_generate(int lattice, const double *box, double radius, int eulerflag, const double *angles, double *out, int cap) {
    double R[9];
    rmatrix(eulerflag, angles, R);
    std::vector<double> v = generate(lattice, box, radius, R);
    int n = (int)(v.size() / 3);
    if (out && cap >= n) std::memcpy(out, v.data(), v.size() * sizeof(double));
    return n;
}

void ora_rmatrix(int flag, const double *angles, double *R) { rmatrix(flag, angles, R); }

void *ora_create(int lattice, double radius, int ptype, int n, const double *xyz, const double *box, int bruteforce) {
    Sys *s = new Sys();
    init_particles(*s, lattice, radius, ptype, n, xyz, box);
    if (bruteforce) create_bonds_bruteforce(*s); else create_bonds(*s);
    if (build_opp(*s) != 0) { delete s; return nullptr; }
    update_connections(*s);
    return s;
}

// Assembly(dumpFile, bondFile, ...) semantics (assembly.h:88-95, 388-424): bonds in file order
void *ora_create_with_bonds(int lattice, double radius, int ptype, int n, const double *xyz, const double *box,
                            int nbonds, const int *blayer, const int *bp1, const int *bp2) {
    Sys *s = new Sys();
    init_particles(*s, lattice, radius, ptype, n, xyz, box);
    std::vector<std::vector<int>> l0(n), l1(n);
    for (int b = 0; b < nbonds; ++b) (blayer[b] == 0 ? l0 : l1)[bp1[b]].push_back(bp2[b]);
    s->bptr.assign(n + 1, 0); s->bsplit.assign(n, 0);
    for (int i = 0; i < n; ++i) s->bptr[i + 1] = s->bptr[i] + (int)(l0[i].size() + l1[i].size());
    s->p2.assign(s->bptr[n], 0); s->layer.assign(s->bptr[n], 0);
    alloc_bond_arrays(*s);
    for (int i = 0; i < n; ++i) {
        int b = s->bptr[i];
        for (int j : l0[i]) { s->p2[b] = j; s->layer[b] = 0; bond_init(*s, b, i, j, distanceTo(&s->xyz[3 * i], &s->xyz[3 * j])); ++b; }
        s->bsplit[i] = b;
        for (int j : l1[i]) { s->p2[b] = j; s->layer[b] = 1; bond_init(*s, b, i, j, distanceTo(&s->xyz[3 * i], &s->xyz[3 * j])); ++b; }
        s->nb[i] = s->bptr[i + 1] - s->bptr[i];
    }
    if (build_opp(*s) != 0) { delete s; return nullptr; }
    update_connections(*s);
    return s;
}

void ora_destroy(void *p) { delete (Sys *)p; }

int ora_info(void *p, const char *what) {
    Sys &s = *(Sys *)p;
    std::string w(what);
    if (w == "N") return s.N;
    if (w == "dim") return s.dim;
    if (w == "nbonds") return (int)s.p2.size();
    if (w == "nconn") return (int)s.cidx.size();
    if (w == "nnl") return (int)s.nlidx.size();
    if (w == "knnz") return (int)s.kcol.size();
    if (w == "nreaction") return (int)s.reaction.size();
    if (w == "cg_iters") return s.cg_iters;
    if (w == "cg_success") return s.cg_success;
    if (w == "newton") return s.last_newton;
    if (w == "n_steps") return s.n_steps_done;
    if (w == "n_restarts") return s.n_restarts;
    if (w == "nsv") return s.nsv;
    if (w == "nneighbors") return s.cell.nneighbors;
    if (w == "newton_log_len") return (int)s.newton_log.size();
    if (w == "cg_log_len") return (int)s.cg_log.size();
    return -1;
}

static std::vector<double> *dfield(Sys &s, const std::string &w) {
    if (w == "xyz") return &s.xyz; if (w == "xyz0") return &s.xyz0; if (w == "xyz_last") return &s.xyz_last;
    if (w == "Pin") return &s.Pin; if (w == "Pex") return &s.Pex; if (w == "stress") return &s.stress;
    if (w == "damage") return &s.damage; if (w == "damage_last") return &s.damage_last;
    if (w == "damage_visual") return &s.damage_visual; if (w == "Ddot_local") return &s.Ddot_local;
    if (w == "Ddot_nonlocal") return &s.Ddot_nonlocal; if (w == "nonlocal_L") return &s.nonlocal_L;
    if (w == "state_var") return &s.sv; if (w == "state_var_last") return &s.sv_last; if (w == "mat") return &s.mat;
    if (w == "dLe_total") return &s.dLe_total; if (w == "TdLe_total") return &s.TdLe_total; if (w == "cs_sum") return &s.cs_sum;
    if (w == "dis0") return &s.dis0; if (w == "dis") return &s.dis; if (w == "dis_last") return &s.dis_last;
    if (w == "dL") return &s.dL; if (w == "dLe") return &s.dLe; if (w == "dLp") return &s.dLp; if (w == "dLp_last") return &s.dLp_last;
    if (w == "bstrain") return &s.bstrain; if (w == "Kn") return &s.Kn; if (w == "Tv") return &s.Tv; if (w == "cs") return &s.cs;
    if (w == "bforce") return &s.bforce; if (w == "bforce_last") return &s.bforce_last;
    if (w == "bdamage") return &s.bdamage; if (w == "bdamage_last") return &s.bdamage_last;
    if (w == "residual") return &s.residual; if (w == "disp") return &s.disp; if (w == "reaction") return &s.reaction;
    if (w == "kval") return &s.kval;
    return nullptr;
}
static std::vector<int> *ifield(Sys &s, const std::string &w) {
    if (w == "bptr") return &s.bptr; if (w == "bsplit") return &s.bsplit; if (w == "p2") return &s.p2;
    if (w == "layer") return &s.layer; if (w == "opp") return &s.opp; if (w == "nb") return &s.nb;
    if (w == "cptr") return &s.cptr; if (w == "cidx") return &s.cidx; if (w == "nconn_largeq") return &s.nconn_largeq;
    if (w == "nlptr") return &s.nlptr; if (w == "nlidx") return &s.nlidx; if (w == "type") return &s.type;
    if (w == "cons") return &s.cons; if (w == "krow") return &s.krow; if (w == "kcol") return &s.kcol;
    if (w == "newton_log") return &s.newton_log; if (w == "cg_log") return &s.cg_log;
    return nullptr;
}

long ora_len(void *p, const char *what) {
    Sys &s = *(Sys *)p;
    if (auto *d = dfield(s, what)) return (long)d->size();
    if (auto *i = ifield(s, what)) return (long)i->size();
    return -1;
}
int ora_get_d(void *p, const char *what, double *out) {
    auto *d = dfield(*(Sys *)p, what);
    if (!d) return -1;
    std::memcpy(out, d->data(), d->size() * sizeof(double));
    return 0;
}
int ora_set_d(void *p, const char *what, const double *in) {
    auto *d = dfield(*(Sys *)p, what);
    if (!d) return -1;
    std::memcpy(d->data(), in, d->size() * sizeof(double));
    return 0;
}
int ora_get_i(void *p, const char *what, int *out) {
    auto *v = ifield(*(Sys *)p, what);
    if (!v) return -1;
    std::memcpy(out, v->data(), v->size() * sizeof(int));
    return 0;
}
int ora_set_i(void *p, const char *what, const int *in) {
    auto *v = ifield(*(Sys *)p, what);
    if (!v) return -1;
    std::memcpy(v->data(), in, v->size() * sizeof(int));
    return 0;
}

int ora_set_mode(void *p, const char *what, double v) {
    Sys &s = *(Sys *)p;
    std::string w(what);
    if (w == "elmap_literal") s.elmap_literal = (int)v;
    else if (w == "diag10_literal") s.diag10_literal = (int)v;
    else if (w == "cg_transpose") s.cg_transpose = (int)v;
    else if (w == "dirichlet_static") s.dirichlet_static = (int)v;
    else if (w == "rollback_pex") s.rollback_pex = (int)v;
    else if (w == "stiffness_fd") s.stiffness_fd = (int)v;
    else if (w == "fd_exact") s.fd_exact = (int)v;
    else if (w == "cg_forcing") s.cg_forcing = (int)v;
    else if (w == "cg_tol") s.cg_tol = v;
    else if (w == "cg_maxit") s.cg_maxit = (long)v;
    else if (w == "max_NR_iter") s.max_NR_iter = (int)v;
    else if (w == "tol_NR") s.tol_NR = v;
    else if (w == "verbose") s.verbose = (int)v;
    else return -1;
    return 0;
}

// setParticleProperty overloads (particle_elastic.h:70-143; particle_elastic_damage.h:175-254;
// particle_j2plasticity.h:154-201; particle_fatigue_hcf.h:162-209). id < 0 applies to all.
static void each(Sys &s, int id, const std::function<void(int)> &f) {
    if (id >= 0) f(id);
    else for (int i = 0; i < s.N; ++i) f(i);
}
void ora_set_elastic(void *p, int id, double L, int ps, double E, double mu, double crit) {
    Sys &s = *(Sys *)p;
    double Ce[3]; ce_from_E(ps != 0, E, mu, Ce);
    each(s, id, [&](int i) { s.mat[10 * i] = crit; s.nonlocal_L[i] = L; set_kntv(s, i, Ce); });
}
void ora_set_elastic_c(void *p, int id, double L, double C11, double C12, double C44, double crit) {
    Sys &s = *(Sys *)p;
    double Ce[3] = {C11, C12, C44};
    each(s, id, [&](int i) { s.mat[10 * i] = crit; s.nonlocal_L[i] = L; set_kntv(s, i, Ce); });
}
void ora_set_elastic_damage(void *p, int id, double L, int ps, double E, double mu, double k0, double k1, double ct, double dthr) {
    Sys &s = *(Sys *)p;
    double Ce[3]; ce_from_E(ps != 0, E, mu, Ce);
    each(s, id, [&](int i) {
        double *m = &s.mat[10 * i]; m[0] = k0; m[1] = k1; m[2] = dthr; m[3] = ct;
        s.nonlocal_L[i] = L; set_kntv(s, i, Ce);
    });
}
void ora_set_j2(void *p, int id, double L, int ps, double E, double mu, double sy, double xi, double H, double A, double crit) {
    Sys &s = *(Sys *)p;
    double Ce[3]; ce_from_E(ps != 0, E, mu, Ce);
    each(s, id, [&](int i) {
        double *m = &s.mat[10 * i]; m[0] = E; m[1] = mu; m[2] = sy; m[3] = xi; m[4] = H; m[5] = A; m[7] = crit;
        s.nonlocal_L[i] = L; set_kntv(s, i, Ce);
    });
}
// damage_threshold of J2 is never set by setParticleProperty (particle_j2plasticity.h:22,154-164)
void ora_set_j2_damage_threshold(void *p, int id, double thr) {
    Sys &s = *(Sys *)p;
    each(s, id, [&](int i) { s.mat[10 * i + 6] = thr; });
}
void ora_set_hcf(void *p, int id, double L, int ps, double E, double mu, double A, double B, double d, double dthr, double flr) {
    Sys &s = *(Sys *)p;
    double Ce[3]; ce_from_E(ps != 0, E, mu, Ce);
    each(s, id, [&](int i) {
        double *m = &s.mat[10 * i]; m[0] = A; m[1] = B; m[2] = d; m[3] = E; m[4] = mu; m[5] = dthr; m[6] = flr;
        s.nonlocal_L[i] = L; set_kntv(s, i, Ce);
    });
}

void ora_search_nonlocal(void *p, double ratio) { search_nonlocal(*(Sys *)p, ratio); }

// isValid over particles (the tools' filter, geometry/geo1_CT_configuration.cpp:194-196)
void ora_cut_particles(int n, const double *xyz, int ncut, const int *kind, const double *par, unsigned char *keep) {
    for (int i = 0; i < n; ++i) keep[i] = is_valid(&xyz[3 * i], ncut, kind, par) ? 1 : 0;
}
// searchNeighbor (geometry/geo1_CT_configuration.cpp:79-134): all pairs, bonds within the cutoff
// whose 10 sub-segment midpoints are all valid, listed per particle i in increasing j as
// (layer, i, j). Returns the count; fills the arrays if cap allows.
int ora_search_neighbor_cut(int lattice, double radius, int n, const double *xyz, int ncut, const int *kind,
                            const double *par, int cap, int *layer, int *p1, int *p2) {
    const Cell cell = make_cell(lattice, radius);
    std::vector<std::vector<int>> nbr(n), lay(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            const double *a = &xyz[3 * i], *b = &xyz[3 * j];
            double dis = std::sqrt(std::pow((b[0] - a[0]), 2) + std::pow((b[1] - a[1]), 2) + std::pow((b[2] - a[2]), 2));
            if ((dis < 1.01 * cell.cutoff[1]) && (j != i)) {
                bool valid = true;
                int n_interval = 10;
                double dx = (b[0] - a[0]) / n_interval;
                double dy = (b[1] - a[1]) / n_interval;
                double dz = (b[2] - a[2]) / n_interval;
                for (int k = 0; k < n_interval; ++k) {
                    double pt1[3] = {a[0] + k * dx, a[1] + k * dy, a[2] + k * dz};
                    double pt2[3] = {a[0] + (k + 1) * dx, a[1] + (k + 1) * dy, a[2] + (k + 1) * dz};
                    double mid[3] = {(pt1[0] + pt2[0]) / 2.0, (pt1[1] + pt2[1]) / 2.0, (pt1[2] + pt2[2]) / 2.0};
                    valid = is_valid(mid, ncut, kind, par) && valid;
                }
                if (valid) {
                    nbr[i].push_back(j);
                    lay[i].push_back(dis < 1.01 * cell.cutoff[0] ? 0 : 1);
                }
            }
        }
    }
    int m = 0;
    for (int i = 0; i < n; ++i)
        for (size_t k = 0; k < nbr[i].size(); ++k, ++m)
            if (m < cap) {
                layer[m] = lay[i][k];
                p1[m] = i;
                p2[m] = nbr[i][k];
            }
    return m;
}
void ora_update_geometry(void *p) { update_geometry(*(Sys *)p); }
void ora_update_force_state(void *p) { update_force_state(*(Sys *)p); }
void ora_update_rr(void *p) { update_rr(*(Sys *)p); }
double ora_norm(void *p, const char *what) {
    auto *d = dfield(*(Sys *)p, what);
    return d ? norm2(*d) : -1.0;
}
void ora_calc_stiffness(void *p) { calc_stiffness(*(Sys *)p); }
void ora_update_stiffness_disp_bc(void *p) { update_stiffness_disp_bc(*(Sys *)p); }
int ora_cg(void *p) { return cg_solve(*(Sys *)p); }
double ora_cg_error(void *p) { return ((Sys *)p)->cg_error; }
void ora_solve_linear_system(void *p) { solve_linear_system(*(Sys *)p); }
int ora_newton(void *p) { return newton_iteration(*(Sys *)p); }
void ora_update_state_var(void *p) { update_state_var(*(Sys *)p); }
int ora_update_static_damage(void *p, int undamaged) { return update_static_damage(*(Sys *)p, undamaged); }
int ora_update_fatigue_damage(void *p, double dNdt, int undamaged) { return update_fatigue_damage(*(Sys *)p, dNdt, undamaged); }
void ora_store(void *p) { store_state(*(Sys *)p); }
void ora_reset(void *p, int reset_xyz) { reset_state(*(Sys *)p, reset_xyz != 0); }
void ora_write_dump(void *p, const char *path, int step) { write_dump(*(Sys *)p, path, step); }

// load steps
void *ora_ls_new() { return new LoadStep(); }
void ora_ls_free(void *l) { delete (LoadStep *)l; }
void ora_ls_add_disp(void *l, int n, const int *ids, int mode, int flag, double step) {
    ((LoadStep *)l)->disp.push_back({std::vector<int>(ids, ids + n), mode, (char)flag, step});
}
void ora_ls_add_force(void *l, int n, const int *ids, int mode, double fx, double fy, double fz) {
    ((LoadStep *)l)->force.push_back({std::vector<int>(ids, ids + n), mode, fx, fy, fz});
}
void ora_apply_bc(void *p, void *l) { apply_force_bc(*(Sys *)p, *(LoadStep *)l); apply_disp_bc(*(Sys *)p, *(LoadStep *)l); }
int ora_solve_step_static(void *p, void *l, int undamaged) { return solve_step_static(*(Sys *)p, *(LoadStep *)l, undamaged); }
int ora_solve_step_fatigue(void *p, void *l) { return solve_step_fatigue(*(Sys *)p, *(LoadStep *)l); }
void ora_solve_problem_static(void *p, int nls, void **ls, int undamaged) {
    std::vector<LoadStep> v;
    for (int k = 0; k < nls; ++k) v.push_back(*(LoadStep *)ls[k]);
    solve_problem_static(*(Sys *)p, v, undamaged);
}
void ora_solve_problem_fatigue_static(void *p, int nls, void **ls) {
    std::vector<LoadStep> v;
    for (int k = 0; k < nls; ++k) v.push_back(*(LoadStep *)ls[k]);
    solve_problem_fatigue_static(*(Sys *)p, v);
}
int ora_solve_problem_cyclic(void *p, int ltype, int n_interval, int nspec, const double *spec, int ngroups,
                             const int *gsizes, const int *gids, int tmode, double tau, int undamaged, int max_cycles_run) {
    std::vector<std::vector<int>> g;
    int off = 0;
    for (int k = 0; k < ngroups; ++k) { g.emplace_back(gids + off, gids + off + gsizes[k]); off += gsizes[k]; }
    std::vector<double> sp(spec, spec + nspec);
    return solve_problem_cyclic(*(Sys *)p, ltype, n_interval, sp, g, tmode, tau, undamaged, max_cycles_run);
}
// one local stiffness block (analytic or FD per mode), row-major 3x3
void ora_local_stiffness(void *p, int i, int j, double *out) {
    Sys &s = *(Sys *)p;
    Blk K = s.stiffness_fd ? local_stiffness_fd(s, i, j) : local_stiffness_ana(s, i, j);
    for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) out[3 * r + q] = K[r][q];
}
double ora_func_phi(double x, double L) { return func_phi(x, L); }
double ora_strain_energy(void *p, int i) { return strain_energy(*(Sys *)p, i); }

}  // extern "C"
