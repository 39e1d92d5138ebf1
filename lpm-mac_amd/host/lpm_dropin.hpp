// C++ drop-in for the reference's header-only API (longfish/LPM-MAC include/*.h).
//
// The class names, public members and method signatures follow the reference so that
// its driver programs (examples/*.cpp: build a lattice, an Assembly, set particle
// properties, define LoadSteps, run SolverStatic / SolverFatigue) compile against these
// headers. The per-step arithmetic runs on the GPU through the C ABI of
// liblpm_mi355x.so (include/lpm_mi355x.h); the Particle / Bond objects here are a host
// mirror of the device state, pushed before device work and pulled back after it.
//
// Differences a caller can observe (each one deliberate):
//  - Solver modes default to the well-posed formulation (Dirichlet DoFs eliminated in the
//    static solver too, symmetric K handed to CG, FP32 CG with FP64 refinement).
//    lpm::options() selects the reference's literal behaviour switch by switch.
//  - Particle ids of each Assembly start at 0 (the reference uses one global counter, so it
//    supports a single Assembly per process).
//  - Bond ids are assigned in bond-list order (the reference's are racy under OpenMP).
#pragma once

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lpm_mi355x.h"

#if __has_include(<Eigen/Dense>)
#include <Eigen/Dense>
#define LPM_HAVE_EIGEN 1
#endif

// ------------------------------------------------------------------ lpm.h
#ifndef PI
#define PI 3.14159265358979323846
#endif
#ifndef EPS
#define EPS 1e-6
#endif
#ifndef NDIM
#define NDIM 3
#endif

// include/lpm.h:45-49
inline double func_phi(const double x, const double L) {
    double p = 8, q = 2;
    return std::pow(1 / (1 + std::pow(x / L, p)), q);
}

enum class ExactSolution : char { Beam001 };
enum class LoadMode : char { Relative, Absolute };
enum class SolverMode : char { CG };
enum class StiffnessMode : char { Analytical, FiniteDifference };
enum class TimeMapMode : char { Linear, Exponential };
enum class ParticleType : char { Elastic, ElasticDamage, J2Plasticity, FatigueHCF };
enum class LatticeType : char { Square2D, Hexagon2D, SimpleCubic3D, FCC3D, BCC3D_1Slip, BCC3D_2Slip, BCC3D_3Slip };

namespace lpm {

// 3x3 matrix used where the reference takes an Eigen::MatrixXd (rotation, el_mapping)
struct Mat3 {
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double &operator()(int r, int c) { return a[3 * r + c]; }
    double operator()(int r, int c) const { return a[3 * r + c]; }
    int rows() const { return 3; }
    int cols() const { return 3; }
};

#ifdef LPM_HAVE_EIGEN
using RMatrix = Eigen::MatrixXd;
inline Mat3 to_mat3(const Eigen::MatrixXd &m) {
    Mat3 o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o(r, c) = m(r, c);
    return o;
}
#else
using RMatrix = Mat3;
inline Mat3 to_mat3(const Mat3 &m) { return m; }
#endif

inline int lattice_id(LatticeType t) {
    switch (t) {
        case LatticeType::Square2D: return LPM_SQUARE2D;
        case LatticeType::Hexagon2D: return LPM_HEXAGON2D;
        case LatticeType::SimpleCubic3D: return LPM_SC3D;
        case LatticeType::FCC3D: return LPM_FCC3D;
        default: return LPM_BCC3D;
    }
}

inline void check(int rc, lpm_ctx *c, const char *what) {
    if (rc != 0) {
        std::string m = std::string(what) + ": " + lpm_last_error(c);
        throw std::runtime_error(m);
    }
}

// Solver switches (lpm_set_mode names); defaults are the well-posed formulation.
struct Options {
    int dirichlet_static = 1;  // reference: 0 (solver_static.h:113)
    int cg_transpose = 0;      // reference: 1 (Eigen CG<Lower|Upper> on column-major K)
    int diag10_literal = 0;    // reference: 1 (derivative.h:38)
    int rollback_pex = 0;      // reference: 0
    int precision = 0;         // 0 mixed FP32/FP64, 1 FP64
    int precond = 0;           // 0 Jacobi, 1 block Jacobi, 2 Jacobi + rigid-body coarse space, 3 block Jacobi + coarse
    int elmap_literal = 1;     // reference: 1 (column-major read of el_mapping)
    double cg_tol = 1e-6;      // reference: 2.22e-16 (FP64)
    double cg_maxit = -1;      // reference: 2*ndof
    int cg_forcing = 0;        // 1: inexact-Newton CG tolerance per Newton iteration (reference: 0)
    int predictor = 0;         // 1: static steps start from the previous converged increment (reference: 0)
    int coarse_reuse = 0;      // 1: keep the two-level coarse space across assemblies while it still works
    int async_dump = 1;        // 1: writeDump from a device snapshot on a writer thread (same text; reference: inline)
    int verbose = 1;           // print like the reference
    int device = 0;
};
inline Options &options() {
    static Options o;
    return o;
}

}  // namespace lpm

#ifndef LPM_HAVE_EIGEN
#ifndef LPM_NO_EIGEN_ALIAS
// Drivers written against the reference spell the rotation type Eigen::MatrixXd
// (e.g. examples/ex1_isoelasticity_2d_hex.cpp:22). Without Eigen installed this alias
// keeps them compiling unchanged; define LPM_NO_EIGEN_ALIAS to turn it off.
namespace Eigen {
using MatrixXd = lpm::Mat3;
}
#endif
#endif

// ------------------------------------------------------------------ unit_cell.h
// UnitCell (unit_cell.h:7-127); constants from lpm_unit_cell
class UnitCell {
   public:
    LatticeType lattice;
    int dim{0};
    int nneighbors{0}, nneighbors1{0}, nneighbors2{0};
    int nneighbors_AFEM{0}, nneighbors_AFEM1{0}, nneighbors_AFEM2{0};
    double radius{0};
    double particle_volume{0};
    lpm::Mat3 el_mapping;  // as the reference reads it (column-major Map, unit_cell.h:49)
    std::vector<double> neighbor_cutoff;

    UnitCell(LatticeType p_lattice, double p_radius) : lattice(p_lattice), radius(p_radius) {
        double out[8], map[9];
        lpm::check(lpm_unit_cell(lpm::lattice_id(lattice), radius, 1, out, map), nullptr, "UnitCell");
        dim = (int)out[0];
        nneighbors = (int)out[1];
        nneighbors1 = (int)out[2];
        nneighbors2 = (int)out[3];
        nneighbors_AFEM = (int)out[4];
        neighbor_cutoff = {out[5], out[6]};
        particle_volume = out[7];
        for (int k = 0; k < 9; ++k) el_mapping.a[k] = map[k];
        // nneighbors_AFEM1/2 (unit_cell.h:56-57,72-73,90-91,107-108,124-125)
        switch (lattice) {
            case LatticeType::Square2D: nneighbors_AFEM1 = 12; nneighbors_AFEM2 = 12; break;
            case LatticeType::Hexagon2D: nneighbors_AFEM1 = 18; nneighbors_AFEM2 = 18; break;
            case LatticeType::SimpleCubic3D: nneighbors_AFEM1 = 24; nneighbors_AFEM2 = 54; break;
            case LatticeType::FCC3D: nneighbors_AFEM1 = 54; nneighbors_AFEM2 = 24; break;
            default: nneighbors_AFEM1 = 34; nneighbors_AFEM2 = 24; break;
        }
    }
};

// ------------------------------------------------------------------ utilities.h
// createRMatrix (utilities.h:8-82)
inline lpm::RMatrix createRMatrix(int eulerflag, double angles[]) {
    double R[9];
    lpm::check(lpm_rotation_matrix(eulerflag, angles, R), nullptr, "createRMatrix");
#ifdef LPM_HAVE_EIGEN
    Eigen::MatrixXd m(3, 3);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m(r, c) = R[3 * r + c];
    return m;
#else
    lpm::Mat3 m;
    for (int k = 0; k < 9; ++k) m.a[k] = R[k];
    return m;
#endif
}

namespace lpm {
inline std::vector<std::array<double, NDIM>> generate(int lattice, const std::array<double, 2 * NDIM> &box, const UnitCell &cell,
                                                      const RMatrix &Rm) {
    Mat3 R = to_mat3(Rm);
    int64_t n = 0;
    check(lpm_generate_lattice_r(lattice, box.data(), cell.radius, R.a, nullptr, 0, &n), nullptr, "lattice generator");
    std::vector<std::array<double, NDIM>> out(n);
    if (n > 0) check(lpm_generate_lattice_r(lattice, box.data(), cell.radius, R.a, out[0].data(), n, &n), nullptr, "lattice generator");
    return out;
}
}  // namespace lpm

// createPlateSQ2D / createPlateHEX2D / createCuboidSC3D / createCuboidFCC3D (utilities.h:84-301)
inline std::vector<std::array<double, NDIM>> createPlateSQ2D(std::array<double, 2 * NDIM> box, UnitCell cell, lpm::RMatrix R) {
    return lpm::generate(LPM_SQUARE2D, box, cell, R);
}
inline std::vector<std::array<double, NDIM>> createPlateHEX2D(std::array<double, 2 * NDIM> box, UnitCell cell, lpm::RMatrix R) {
    return lpm::generate(LPM_HEXAGON2D, box, cell, R);
}
inline std::vector<std::array<double, NDIM>> createCuboidSC3D(std::array<double, 2 * NDIM> box, UnitCell cell, lpm::RMatrix R) {
    return lpm::generate(LPM_SC3D, box, cell, R);
}
inline std::vector<std::array<double, NDIM>> createCuboidFCC3D(std::array<double, 2 * NDIM> box, UnitCell cell, lpm::RMatrix R) {
    return lpm::generate(LPM_FCC3D, box, cell, R);
}

// ------------------------------------------------------------------ bond.h / particle.h
template <int nlayer>
class Particle;

// Bond (bond.h:12-73)
template <int nlayer>
class Bond {
   public:
    int id{0};
    int layer{-1};
    double dis_initial{0}, dis_last{0}, dis{0};
    double dL{0}, dLe{0}, dLp{0}, dLp_last{0};
    double dLe_dis{0};
    double bstrain{0};
    double Kn{0}, Tv{0};
    double csx{0}, csy{0}, csz{0};
    double bforce_last{0}, bforce{0}, bdamage{0}, bdamage_last{0};
    Particle<nlayer> *p1{nullptr}, *p2{nullptr};

    Bond(Particle<nlayer> *a, Particle<nlayer> *b, int p_layer, double p_dis) : layer(p_layer), dis_initial(p_dis), dis_last(p_dis),
                                                                                 dis(p_dis), p1(a), p2(b) {
        csx = (p1->xyz[0] - p2->xyz[0]) / dis;
        csy = (p1->xyz[1] - p2->xyz[1]) / dis;
        csz = (p1->xyz[2] - p2->xyz[2]) / dis;
    }
};

template <int nlayer>
class Assembly;

// Particle (particle.h:13-66). Model parameters set by the derived classes'
// setParticleProperty are kept in ce/params and pushed to the device by the Assembly.
template <int nlayer>
class Particle {
   public:
    int id{0};
    int type{0};
    int nconn_largeq{0};
    int nb{0}, nconn{0};
    double nonlocal_L{0.};
    double damage_last{0.}, damage{0.}, damage_visual{0.};
    double Ddot_local{0}, Ddot_nonlocal{0};
    UnitCell cell;
    std::vector<double> state_var, state_var_last;
    std::array<int, NDIM> disp_constraint{0, 0, 0};
    std::array<double, nlayer> TdLe_total{}, dLe_total{}, cs_sumx{}, cs_sumy{}, cs_sumz{};
    std::array<double, NDIM> xyz{}, xyz_initial{}, xyz_last{};
    std::array<double, NDIM> Pin{0, 0, 0}, Pex{0, 0, 0};
    std::array<std::vector<Bond<nlayer> *>, nlayer> bond_layers;
    std::vector<Particle<nlayer> *> neighbors;
    std::vector<Particle<nlayer> *> conns;
    std::vector<Particle<nlayer> *> neighbors_nonlocal;
    std::vector<double> stress, strain;
    // model parameters (lpm_set_material layout) and the mapping input
    std::array<double, 10> params{};
    std::array<double, 3> ce{0, 0, 0};
    bool has_property{false};

    Particle(const double &x, const double &y, const double &z, const UnitCell &p_cell) : cell(p_cell) {
        xyz = {x, y, z};
        xyz_initial = xyz;
        xyz_last = xyz;
        stress.assign(2 * NDIM, 0.0);
    }
    virtual ~Particle() {}

    void moveTo(const double &x, const double &y, const double &z) { xyz = {x, y, z}; }
    void moveTo(const std::array<double, NDIM> &p) { xyz = p; }
    void moveBy(const std::array<double, NDIM> &d) { xyz = {xyz[0] + d[0], xyz[1] + d[1], xyz[2] + d[2]}; }
    double distanceTo(Particle<nlayer> *A) {
        return std::sqrt(std::pow((xyz[0] - A->xyz[0]), 2) + std::pow((xyz[1] - A->xyz[1]), 2) + std::pow((xyz[2] - A->xyz[2]), 2));
    }
    bool operator==(const Particle<nlayer> &o) { return id == o.id; }
    // particle.h:242-256
    bool hasAFEMneighbor(Particle<nlayer> *pj, int layer) {
        for (Bond<nlayer> *b1 : bond_layers[layer]) {
            if (b1->p2->id == pj->id) return true;
            for (Bond<nlayer> *b2 : b1->p2->bond_layers[layer])
                if (b2->p2->id == pj->id) return true;
        }
        return false;
    }
    // KnTv = el_mapping * Ce (x radius in 3D) and per-bond Kn/Tv (particle_elastic.h:90-110)
    void applyCe(double L, const std::array<double, 3> &Ce) {
        nonlocal_L = L;
        ce = Ce;
        has_property = true;
        const lpm::Mat3 M = lpm::options().elmap_literal ? cell.el_mapping : rowMajorMap();
        double KnTv[3];
        for (int r = 0; r < 3; ++r) {
            double s = M(r, 0) * Ce[0];
            s = s + M(r, 1) * Ce[1];
            s = s + M(r, 2) * Ce[2];
            KnTv[r] = cell.dim == 2 ? s : s * cell.radius;
        }
        for (int l = 0; l < nlayer; ++l)
            for (Bond<nlayer> *bd : bond_layers[l]) {
                if (cell.lattice == LatticeType::Hexagon2D) { bd->Kn = KnTv[0]; bd->Tv = KnTv[1]; }
                else { bd->Kn = KnTv[bd->layer]; bd->Tv = KnTv[2]; }
            }
    }
    lpm::Mat3 rowMajorMap() const {
        lpm::Mat3 m;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) m(r, c) = cell.el_mapping(c, r);
        return m;
    }
    static std::array<double, 3> ceFromE(bool plane_stress, double E, double mu) {
        std::array<double, 3> c;
        lpm_ce_from_e(plane_stress ? 1 : 0, E, mu, c.data());
        return c;
    }
};

// ParticleElastic (particle_elastic.h:15-143)
template <int nlayer>
class ParticleElastic : public Particle<nlayer> {
   public:
    ParticleElastic(const double &x, const double &y, const double &z, const UnitCell &c) : Particle<nlayer>(x, y, z, c) {
        this->state_var = {0};
        this->state_var_last = {0};
    }
    void setParticleProperty(double L, bool plane_stress, double E, double mu, double crit) {
        this->params = {};
        this->params[0] = crit;
        this->applyCe(L, Particle<nlayer>::ceFromE(plane_stress, E, mu));
    }
    void setParticleProperty(double L, double C11, double C12, double C44, double crit) {
        this->params = {};
        this->params[0] = crit;
        this->applyCe(L, {C11, C12, C44});
    }
};

// ParticleElasticDamage (particle_elastic_damage.h:168-254)
template <int nlayer>
class ParticleElasticDamage : public Particle<nlayer> {
   public:
    double k0{0}, k1{0}, damage_threshold{0}, comp_tensile_ratio{0};
    ParticleElasticDamage(const double &x, const double &y, const double &z, const UnitCell &c) : Particle<nlayer>(x, y, z, c) {
        this->state_var = std::vector<double>(1, 0.);
        this->state_var_last = std::vector<double>(1, 0.);
    }
    void setParticleProperty(double L, bool ps, double E, double mu, double p_k0, double p_k1, double ct, double thr) {
        k0 = p_k0; k1 = p_k1; damage_threshold = thr; comp_tensile_ratio = ct;
        this->params = {k0, k1, thr, ct};
        this->applyCe(L, Particle<nlayer>::ceFromE(ps, E, mu));
    }
    void setParticleProperty(double L, double C11, double C12, double C44, double p_k0, double p_k1, double ct, double thr) {
        k0 = p_k0; k1 = p_k1; damage_threshold = thr; comp_tensile_ratio = ct;
        this->params = {k0, k1, thr, ct};
        this->applyCe(L, {C11, C12, C44});
    }
};

// ParticleJ2Plasticity (particle_j2plasticity.h:17-201); damage_threshold is never set by
// setParticleProperty in the reference (stays 0) -- assign the member to change it.
template <int nlayer>
class ParticleJ2Plasticity : public Particle<nlayer> {
   public:
    double E{0}, mu{0}, sigmay{0}, xi{0}, H{0}, A{0};
    double damage_threshold{0}, critical_bstrain{0};
    ParticleJ2Plasticity(const double &x, const double &y, const double &z, const UnitCell &c) : Particle<nlayer>(x, y, z, c) {
        this->state_var = std::vector<double>(7, 0.);
        this->state_var_last = std::vector<double>(7, 0.);
    }
    void setParticleProperty(double L, bool ps, double pE, double pmu, double psy, double pxi, double pH, double pA, double crit) {
        E = pE; mu = pmu; sigmay = psy; xi = pxi; H = pH; A = pA; critical_bstrain = crit;
        this->params = {E, mu, sigmay, xi, H, A, damage_threshold, crit};
        this->applyCe(L, Particle<nlayer>::ceFromE(ps, E, mu));
    }
};

// ParticleFatigueHCF (particle_fatigue_hcf.h:162-209)
template <int nlayer>
class ParticleFatigueHCF : public Particle<nlayer> {
   public:
    double A{0}, B{0}, d{0};
    double E{0}, mu{0}, fatigue_limit_ratio{1.108};
    double damage_threshold{1};
    ParticleFatigueHCF(const double &x, const double &y, const double &z, const UnitCell &c) : Particle<nlayer>(x, y, z, c) {
        this->state_var = {0, 2147483647.0, 0};
        this->state_var_last = {0, 2147483647.0, 0};
        this->params = {};
        this->params[5] = 1.0;
        this->params[6] = 1.108;
    }
    void setParticleProperty(double L, bool ps, double pE, double pmu, double pA, double pB, double pd, double thr, double flr) {
        E = pE; mu = pmu; A = pA; B = pB; d = pd; damage_threshold = thr; fatigue_limit_ratio = flr;
        this->params = {A, B, d, E, mu, thr, flr};
        this->applyCe(L, Particle<nlayer>::ceFromE(ps, E, mu));
    }
};

// ------------------------------------------------------------------ boundary.h / load_step.h
template <int nlayer>
class DispBC {
   public:
    std::vector<Particle<nlayer> *> group;
    LoadMode load_mode;
    char flag;
    double step;
    DispBC(std::vector<Particle<nlayer> *> g, LoadMode m, char f, double s) : group{g}, load_mode{m}, flag{f}, step{s} {}
    void cutHalf() { step /= 2.0; }
};

template <int nlayer>
class ForceBC {
   public:
    std::vector<Particle<nlayer> *> group;
    LoadMode load_mode;
    double fx, fy, fz;
    ForceBC(std::vector<Particle<nlayer> *> g, LoadMode m, double x, double y, double z) : group{g}, load_mode{m}, fx{x}, fy{y}, fz{z} {}
    void cutHalf() { fx /= 2.0, fy /= 2.0, fz /= 2.0; }
};

template <int nlayer>
class LoadStep {
   public:
    std::vector<DispBC<nlayer>> dispBCs;
    std::vector<ForceBC<nlayer>> forceBCs;
    LoadStep() {}
    void loadCutHalf() {
        for (auto &d : dispBCs) d.cutHalf();
        for (auto &f : forceBCs) f.cutHalf();
    }
};

// load_step_fatigue.h:13-85
enum class FatigueLoadType : char { LoadUniaxialDisp, LoadUniaxialForce, LoadDogBoneDisp, LoadCTForce };

template <int nlayer>
class LoadUniaxialForce : public LoadStep<nlayer> {
   public:
    LoadUniaxialForce(double fy, std::vector<Particle<nlayer> *> &left, std::vector<Particle<nlayer> *> &top,
                      std::vector<Particle<nlayer> *> &bottom) {
        this->dispBCs.push_back(DispBC<nlayer>(left, LoadMode::Relative, 'x', 0.0));
        this->forceBCs.push_back(ForceBC<nlayer>(top, LoadMode::Relative, 0.0, fy, 0.0));
        this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'y', 0.0));
    }
};
template <int nlayer>
class LoadUniaxialDisp : public LoadStep<nlayer> {
   public:
    LoadUniaxialDisp(double dy, std::vector<Particle<nlayer> *> &left, std::vector<Particle<nlayer> *> &top,
                     std::vector<Particle<nlayer> *> &bottom) {
        this->dispBCs.push_back(DispBC<nlayer>(left, LoadMode::Relative, 'x', 0.0));
        this->dispBCs.push_back(DispBC<nlayer>(top, LoadMode::Relative, 'y', dy));
        this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'y', 0.0));
    }
};
template <int nlayer>
class LoadDogBoneDisp : public LoadStep<nlayer> {
   public:
    LoadDogBoneDisp(double x, double y, double z, std::vector<Particle<nlayer> *> &g1, std::vector<Particle<nlayer> *> &g2) {
        this->dispBCs.push_back(DispBC<nlayer>(g1, LoadMode::Relative, 'x', x));
        this->dispBCs.push_back(DispBC<nlayer>(g1, LoadMode::Relative, 'y', y));
        this->dispBCs.push_back(DispBC<nlayer>(g1, LoadMode::Relative, 'z', z));
        this->dispBCs.push_back(DispBC<nlayer>(g2, LoadMode::Relative, 'x', 0.0));
        this->dispBCs.push_back(DispBC<nlayer>(g2, LoadMode::Relative, 'y', 0.0));
        this->dispBCs.push_back(DispBC<nlayer>(g2, LoadMode::Relative, 'z', 0.0));
    }
};
template <int nlayer>
class LoadCTForce : public LoadStep<nlayer> {
   public:
    LoadCTForce(double fy, std::vector<Particle<nlayer> *> &mid, std::vector<Particle<nlayer> *> &top,
                std::vector<Particle<nlayer> *> &bottom) {
        this->dispBCs.push_back(DispBC<nlayer>(mid, LoadMode::Relative, 'y', 0));
        this->dispBCs.push_back(DispBC<nlayer>(top, LoadMode::Relative, 'x', 0.0));
        this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'x', 0.0));
        this->forceBCs.push_back(ForceBC<nlayer>(top, LoadMode::Relative, 0.0, fy, 0.0));
        this->forceBCs.push_back(ForceBC<nlayer>(bottom, LoadMode::Relative, 0.0, -fy, 0.0));
    }
};

// ------------------------------------------------------------------ assembly.h
namespace lpm {
// Device side of one particle system, shared by all copies of an Assembly (the reference's
// Solver keeps a copy of the Assembly that shares its particles, solver.h:29,45).
struct Device {
    lpm_ctx *ctx{nullptr};
    std::map<std::vector<int32_t>, int32_t> groups;
    ~Device() {
        if (ctx) lpm_destroy(ctx);
    }
};
}  // namespace lpm

template <int nlayer>
class Assembly {
   public:
    ParticleType ptype;
    int nparticle{0};
    std::vector<Particle<nlayer> *> pt_sys;
    std::array<double, 2 * NDIM> box{};
    std::shared_ptr<lpm::Device> dev;

    // assembly.h:97-106
    Assembly(std::vector<std::array<double, NDIM>> &p_xyz, std::array<double, 2 * NDIM> &p_box, UnitCell &p_cell,
             const ParticleType &p_ptype)
        : ptype(p_ptype), box(p_box) {
        createParticles(p_xyz, p_cell);
        createBonds();
        updateConnections();
        nparticle = (int)pt_sys.size();
    }
    // assembly.h:73-80
    Assembly(const std::string &dumpFile, UnitCell &p_cell, const ParticleType &p_ptype) : ptype(p_ptype) {
        printf("Reading data ...\n");
        readDump(dumpFile, p_cell);
        createBonds();
        updateConnections();
    }
    // assembly.h:82-95
    Assembly(const std::string &dumpFile, const std::string &bondFile, UnitCell &p_cell, const ParticleType &p_ptype)
        : ptype(p_ptype) {
        printf("Reading data ...\n");
        readDump(dumpFile, p_cell);
        readBond(bondFile);
        for (Particle<nlayer> *p1 : pt_sys) p1->nb = (int)p1->neighbors.size();
        updateConnections();
    }

    std::map<int, Particle<nlayer> *> toMap() {
        std::map<int, Particle<nlayer> *> m;
        for (auto *pt : pt_sys) m[pt->id] = pt;
        return m;
    }

    void createParticles(std::vector<std::array<double, NDIM>> &p_xyz, UnitCell &p_cell) {
        int id = 0;
        for (auto &x : p_xyz) {
            Particle<nlayer> *pt = make(x[0], x[1], x[2], p_cell);
            pt->id = id++;
            pt_sys.push_back(pt);
        }
    }

    // Assembly::createBonds (assembly.h:212-234) on the cell-list search of the library
    void createBonds() {
        const int64_t n = (int64_t)pt_sys.size();
        std::vector<double> xyz(3 * n);
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) xyz[3 * i + k] = pt_sys[i]->xyz[k];
        int64_t nb = 0;
        const UnitCell &c = pt_sys[0]->cell;
        const int lat = lpm::lattice_id(c.lattice);
        lpm::check(lpm_build_bonds(lat, c.radius, n, xyz.data(), nullptr, nullptr, nullptr, 0, &nb), nullptr, "createBonds");
        std::vector<int32_t> bp(n + 1), bs(n), p2(nb);
        lpm::check(lpm_build_bonds(lat, c.radius, n, xyz.data(), bp.data(), bs.data(), p2.data(), nb, &nb), nullptr, "createBonds");
        int bid = 0;
        for (int64_t i = 0; i < n; ++i) {
            Particle<nlayer> *p1 = pt_sys[i];
            for (int b = bp[i]; b < bp[i + 1]; ++b) {
                Particle<nlayer> *p2p = pt_sys[p2[b]];
                const int layer = b < bs[i] ? 0 : 1;
                Bond<nlayer> *bd = new Bond<nlayer>(p1, p2p, layer, p1->distanceTo(p2p));
                bd->id = bid++;
                p1->bond_layers[layer].push_back(bd);
                p1->neighbors.push_back(p2p);
            }
            p1->nb = (int)p1->neighbors.size();
        }
    }

    // Assembly::updateConnections (assembly.h:236-266): the pattern is built on the device
    void updateConnections() {
        ensureDevice();
        const int64_t n = (int64_t)pt_sys.size();
        std::vector<int32_t> cp(n + 1), ci(lpm_size(dev->ctx, "nconn"));
        lpm::check(lpm_download(dev->ctx, "conn_ptr", cp.data(), n + 1), dev->ctx, "conn_ptr");
        lpm::check(lpm_download(dev->ctx, "conn_idx", ci.data(), (int64_t)ci.size()), dev->ctx, "conn_idx");
        for (int64_t i = 0; i < n; ++i) {
            Particle<nlayer> *p = pt_sys[i];
            p->conns.clear();
            for (int k = cp[i]; k < cp[i + 1]; ++k) p->conns.push_back(pt_sys[ci[k]]);
            p->nconn = (int)p->conns.size();
            auto it = std::find(p->conns.begin(), p->conns.end(), p);
            p->nconn_largeq = (int)std::distance(it, p->conns.end());
        }
    }

    // Assembly::searchNonlocalNeighbors (assembly.h:117-130), on the device
    void searchNonlocalNeighbors(double cutoff_ratio) {
        push();
        lpm::check(lpm_search_nonlocal(dev->ctx, cutoff_ratio), dev->ctx, "searchNonlocalNeighbors");
        std::vector<int32_t> p(pt_sys.size() + 1), q(lpm_size(dev->ctx, "nnl"));
        lpm::check(lpm_download(dev->ctx, "nl_ptr", p.data(), (int64_t)p.size()), dev->ctx, "nl_ptr");
        lpm::check(lpm_download(dev->ctx, "nl_idx", q.data(), (int64_t)q.size()), dev->ctx, "nl_idx");
        for (size_t i = 0; i < pt_sys.size(); ++i) {
            pt_sys[i]->neighbors_nonlocal.clear();
            for (int k = p[i]; k < p[i + 1]; ++k) pt_sys[i]->neighbors_nonlocal.push_back(pt_sys[q[k]]);
        }
    }

    void updateGeometry() { run([&] { return lpm_update_geometry(dev->ctx); }, "updateGeometry"); }
    void updateForceState() { run([&] { return lpm_update_force_state(dev->ctx); }, "updateForceState"); }
    void updateStateVar() { run([&] { return lpm_update_state_var(dev->ctx); }, "updateStateVar"); }
    void storeStateVar() { run([&] { return lpm_store_state(dev->ctx); }, "storeStateVar"); }
    void resetStateVar(bool reset_xyz) { run([&] { return lpm_reset_state(dev->ctx, reset_xyz ? 1 : 0); }, "resetStateVar"); }

    // Assembly::writeBond (assembly.h:268-282)
    void writeBond(const std::string &bondFile) {
        FILE *f = fopen(bondFile.c_str(), "w+");
        if (!f) return;
        for (auto *pt : pt_sys)
            for (int i = 0; i < nlayer; ++i)
                for (auto *bd : pt->bond_layers[i]) fprintf(f, "%d %d %d %d\n", bd->id, bd->layer, bd->p1->id, bd->p2->id);
        fclose(f);
    }
    // Assembly::writeConfigurationDump (assembly.h:284-307)
    void writeConfigurationDump(const std::string &dumpFile) {
        FILE *f = fopen(dumpFile.c_str(), "w+");
        if (!f) return;
        fprintf(f, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n", 0, nparticle);
        fprintf(f, "%8.8f %8.8f\n%8.8f %8.8f\n%8.8f %8.8f\n", box[0], box[1], box[2], box[3], box[4], box[5]);
        fprintf(f, "ITEM: ATOMS id type x y z\n");
        for (auto *pt : pt_sys) fprintf(f, "%d %d %.4e %.4e %.4e \n", pt->id, pt->type, pt->xyz[0], pt->xyz[1], pt->xyz[2]);
        fclose(f);
    }
    // Assembly::writeDump (assembly.h:309-335)
    void writeDump(const std::string &dumpFile, int step) {
        FILE *f = fopen(dumpFile.c_str(), "a+");
        if (!f) return;
        fprintf(f, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n", step, nparticle);
        fprintf(f, "%8.8f %8.8f\n%8.8f %8.8f\n%8.8f %8.8f\n", box[0], box[1], box[2], box[3], box[4], box[5]);
        fprintf(f, "ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 damage damage_visual state_var[0] \n");
        for (auto *pt : pt_sys)
            fprintf(f, "%d %d %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e\n", pt->id, pt->type,
                    pt->xyz[0], pt->xyz[1], pt->xyz[2], pt->xyz[0] - pt->xyz_initial[0], pt->xyz[1] - pt->xyz_initial[1],
                    pt->xyz[2] - pt->xyz_initial[2], pt->stress[0], pt->stress[1], pt->stress[2], pt->stress[3], pt->stress[4],
                    pt->stress[5], pt->damage, pt->damage_visual, pt->state_var[0]);
        fclose(f);
    }
    // writeDump without stopping the solver: the device snapshot is formatted and appended by a
    // writer thread (lpm_dump_async); waitDump() joins it
    void writeDumpAsync(const std::string &dumpFile, int step) {
        ensureDevice();
        lpm::check(lpm_dump_async(dev->ctx, dumpFile.c_str(), step, box.data()), dev->ctx, "writeDump");
    }
    void waitDump() {
        if (dev && dev->ctx) lpm::check(lpm_dump_wait(dev->ctx), dev->ctx, "writeDump");
    }
    // Assembly::readDump (assembly.h:337-386): id type x y z
    void readDump(const std::string &dumpFile, UnitCell &cell) {
        FILE *f = fopen(dumpFile.c_str(), "r+");
        if (f == NULL) {
            printf("\'%s\' does not exist!\n", dumpFile.c_str());
            exit(1);
        }
        char line[500];
        for (int k = 0; k < 3; ++k)
            if (!fgets(line, 500, f)) break;
        if (fscanf(f, "%d\n", &nparticle) != 1) nparticle = 0;
        if (!fgets(line, 500, f)) line[0] = 0;
        for (int k = 0; k < 3; ++k)
            if (fscanf(f, "%lf %lf\n", &box[2 * k], &box[2 * k + 1]) != 2) break;
        if (!fgets(line, 500, f)) line[0] = 0;
        int id, type;
        double x[3];
        while (fscanf(f, "%d %d %lf %lf %lf", &id, &type, &x[0], &x[1], &x[2]) > 0) {
            Particle<nlayer> *pt = make(x[0], x[1], x[2], cell);
            pt->id = id;
            pt->type = type;
            pt_sys.push_back(pt);
        }
        fclose(f);
        nparticle = (int)pt_sys.size();
        for (size_t i = 0; i < pt_sys.size(); ++i)
            if (pt_sys[i]->id != (int)i) throw std::runtime_error("readDump: particle ids must be 0..n-1 in file order");
    }
    // Assembly::readBond (assembly.h:388-424): "id layer p1 p2" in file order
    void readBond(const std::string &bondFile) {
        FILE *f = fopen(bondFile.c_str(), "r+");
        if (f == NULL) {
            printf("\'%s\' does not exist!\n", bondFile.c_str());
            exit(1);
        }
        auto m = toMap();
        int id, layer, a, b;
        while (fscanf(f, "%d %d %d %d", &id, &layer, &a, &b) == 4) {
            Particle<nlayer> *p1 = m[a], *p2 = m[b];
            Bond<nlayer> *bd = new Bond<nlayer>(p1, p2, layer, p1->distanceTo(p2));
            bd->id = id;
            p1->bond_layers[layer].push_back(bd);
            p1->neighbors.push_back(p2);
        }
        fclose(f);
    }

    // ---------------------------------------------------------- host <-> device mirror
    void ensureDevice() {
        if (dev && dev->ctx) return;
        if (!dev) dev = std::make_shared<lpm::Device>();
        const int64_t n = (int64_t)pt_sys.size();
        std::vector<double> xyz(3 * n);
        std::vector<int32_t> bp(n + 1, 0), bs(n), p2;
        for (int64_t i = 0; i < n; ++i) {
            Particle<nlayer> *p = pt_sys[i];
            for (int k = 0; k < 3; ++k) xyz[3 * i + k] = p->xyz[k];
            for (auto *bd : p->bond_layers[0]) p2.push_back(bd->p2->id);
            bs[i] = (int32_t)p2.size();
            for (auto *bd : p->bond_layers[1]) p2.push_back(bd->p2->id);
            bp[i + 1] = (int32_t)p2.size();
        }
        const UnitCell &c = pt_sys[0]->cell;
        lpm_topology t{};
        t.lattice = lpm::lattice_id(c.lattice);
        t.radius = c.radius;
        t.model = (int)ptype;
        t.n = n;
        t.xyz = xyz.data();
        t.bond_ptr = bp.data();
        t.bond_split = bs.data();
        t.bond_p2 = p2.data();
        t.device = lpm::options().device;
        lpm::check(lpm_create(&t, &dev->ctx), nullptr, "lpm_create");
        applyOptions();
    }

    void applyOptions() {
        const lpm::Options &o = lpm::options();
        lpm_ctx *c = dev->ctx;
        lpm_set_mode(c, "dirichlet_static", o.dirichlet_static);
        lpm_set_mode(c, "cg_transpose", o.cg_transpose);
        lpm_set_mode(c, "diag10_literal", o.diag10_literal);
        lpm_set_mode(c, "rollback_pex", o.rollback_pex);
        lpm_set_mode(c, "precision", o.precision);
        lpm_set_mode(c, "precond", o.precond);
        lpm_set_mode(c, "cg_tol", o.cg_tol);
        lpm_set_mode(c, "cg_maxit", o.cg_maxit);
        lpm_set_mode(c, "cg_forcing", o.cg_forcing);
        lpm_set_mode(c, "predictor", o.predictor);
        lpm_set_mode(c, "coarse_reuse", o.coarse_reuse);
        lpm_set_mode(c, "verbose", o.verbose);
    }

    template <typename T>
    void up(const char *f, const std::vector<T> &v) {
        lpm::check(lpm_upload(dev->ctx, f, v.data(), (int64_t)v.size()), dev->ctx, f);
    }
    template <typename T>
    void down(const char *f, std::vector<T> &v) {
        lpm::check(lpm_download(dev->ctx, f, v.data(), (int64_t)v.size()), dev->ctx, f);
    }

    // host mirror -> device (every field a caller may have changed)
    void push() {
        ensureDevice();
        const size_t n = pt_sys.size();
        std::vector<double> xyz(3 * n), xyz_last(3 * n), pex(3 * n), mat(10 * n), L(n), dmg(n), dmg_last(n), sv(7 * n, 0.0),
            svl(7 * n, 0.0);
        std::vector<int32_t> type(n), cons(3 * n);
        std::vector<double> kn, tv, bd, bdl, dlp, dlpl;
        for (size_t i = 0; i < n; ++i) {
            Particle<nlayer> *p = pt_sys[i];
            for (int k = 0; k < 3; ++k) {
                xyz[3 * i + k] = p->xyz[k];
                xyz_last[3 * i + k] = p->xyz_last[k];
                pex[3 * i + k] = p->Pex[k];
                cons[3 * i + k] = p->disp_constraint[k];
            }
            for (int k = 0; k < 10; ++k) mat[10 * i + k] = p->params[k];
            if (auto *j2 = dynamic_cast<ParticleJ2Plasticity<nlayer> *>(p)) mat[10 * i + 6] = j2->damage_threshold;
            L[i] = p->nonlocal_L;
            dmg[i] = p->damage;
            dmg_last[i] = p->damage_last;
            type[i] = p->type;
            for (size_t k = 0; k < p->state_var.size() && k < 7; ++k) {
                sv[7 * i + k] = p->state_var[k];
                svl[7 * i + k] = p->state_var_last[k];
            }
            for (int l = 0; l < nlayer; ++l)
                for (auto *b : p->bond_layers[l]) {
                    kn.push_back(b->Kn); tv.push_back(b->Tv); bd.push_back(b->bdamage); bdl.push_back(b->bdamage_last);
                    dlp.push_back(b->dLp); dlpl.push_back(b->dLp_last);
                }
        }
        up("xyz", xyz); up("xyz_last", xyz_last); up("Pex", pex); up("disp_constraint", cons); up("type", type);
        up("mat", mat); up("nonlocal_L", L); up("damage", dmg); up("damage_last", dmg_last);
        up("state_var", sv); up("state_var_last", svl);
        up("Kn", kn); up("Tv", tv); up("bdamage", bd); up("bdamage_last", bdl); up("dLp", dlp); up("dLp_last", dlpl);
    }

    // device -> host mirror; bonds optional (they are large)
    void pull(bool bonds = true) {
        const size_t n = pt_sys.size();
        std::vector<double> xyz(3 * n), xyz_last(3 * n), pin(3 * n), pex(3 * n), st(6 * n), dmg(n), dmgl(n), dv(n), dl(n), dn(n),
            sv(7 * n), svl(7 * n), dlt(2 * n), tdlt(2 * n), css(6 * n);
        std::vector<int32_t> cons(3 * n);
        down("xyz", xyz); down("xyz_last", xyz_last); down("Pin", pin); down("Pex", pex); down("stress", st);
        down("damage", dmg); down("damage_last", dmgl); down("damage_visual", dv); down("Ddot_local", dl);
        down("Ddot_nonlocal", dn); down("state_var", sv); down("state_var_last", svl); down("dLe_total", dlt);
        down("TdLe_total", tdlt); down("cs_sum", css); down("disp_constraint", cons);
        for (size_t i = 0; i < n; ++i) {
            Particle<nlayer> *p = pt_sys[i];
            for (int k = 0; k < 3; ++k) {
                p->xyz[k] = xyz[3 * i + k]; p->xyz_last[k] = xyz_last[3 * i + k]; p->Pin[k] = pin[3 * i + k];
                p->Pex[k] = pex[3 * i + k]; p->disp_constraint[k] = cons[3 * i + k];
            }
            p->stress.assign(st.begin() + 6 * i, st.begin() + 6 * i + 6);
            p->damage = dmg[i]; p->damage_last = dmgl[i]; p->damage_visual = dv[i];
            p->Ddot_local = dl[i]; p->Ddot_nonlocal = dn[i];
            for (size_t k = 0; k < p->state_var.size() && k < 7; ++k) {
                p->state_var[k] = sv[7 * i + k];
                p->state_var_last[k] = svl[7 * i + k];
            }
            for (int l = 0; l < nlayer; ++l) {
                p->dLe_total[l] = dlt[2 * i + l]; p->TdLe_total[l] = tdlt[2 * i + l];
                p->cs_sumx[l] = css[6 * i + 3 * l]; p->cs_sumy[l] = css[6 * i + 3 * l + 1]; p->cs_sumz[l] = css[6 * i + 3 * l + 2];
            }
        }
        if (!bonds) return;
        const int64_t nb = lpm_size(dev->ctx, "nbonds");
        std::vector<double> dis(nb), disl(nb), dle(nb), dlp(nb), dlpl(nb), bs(nb), cs(3 * nb), bf(nb), bfl(nb), bd(nb), bdl(nb);
        down("dis", dis); down("dis_last", disl); down("dLe", dle); down("dLp", dlp); down("dLp_last", dlpl);
        down("bstrain", bs); down("cs", cs); down("bforce", bf); down("bforce_last", bfl); down("bdamage", bd);
        down("bdamage_last", bdl);
        int64_t b = 0;
        for (auto *p : pt_sys)
            for (int l = 0; l < nlayer; ++l)
                for (auto *x : p->bond_layers[l]) {
                    x->dis = dis[b]; x->dis_last = disl[b]; x->dLe = dle[b]; x->dLp = dlp[b]; x->dLp_last = dlpl[b];
                    x->bstrain = bs[b]; x->dL = x->bstrain * x->dis;
                    x->csx = cs[3 * b]; x->csy = cs[3 * b + 1]; x->csz = cs[3 * b + 2];
                    x->bforce = bf[b]; x->bforce_last = bfl[b]; x->bdamage = bd[b]; x->bdamage_last = bdl[b];
                    ++b;
                }
    }

    int32_t groupId(const std::vector<Particle<nlayer> *> &g) {
        std::vector<int32_t> ids;
        for (auto *p : g) ids.push_back(p->id);
        auto it = dev->groups.find(ids);
        if (it != dev->groups.end()) return it->second;
        int32_t gid = 0;
        lpm::check(lpm_group_create(dev->ctx, (int64_t)ids.size(), ids.data(), &gid), dev->ctx, "group");
        dev->groups[ids] = gid;
        return gid;
    }

   private:
    Particle<nlayer> *make(double x, double y, double z, UnitCell &cell) {
        switch (ptype) {
            case ParticleType::Elastic: return new ParticleElastic<nlayer>(x, y, z, cell);
            case ParticleType::ElasticDamage: return new ParticleElasticDamage<nlayer>(x, y, z, cell);
            case ParticleType::J2Plasticity: return new ParticleJ2Plasticity<nlayer>(x, y, z, cell);
            default: return new ParticleFatigueHCF<nlayer>(x, y, z, cell);
        }
    }
    template <typename F>
    void run(F f, const char *what) {
        push();
        lpm::check(f(), dev->ctx, what);
        pull();
    }
};

// ------------------------------------------------------------------ stiffness.h
// Stiffness (stiffness.h:18-41): the matrix lives on the device in BSR form.
template <int nlayer>
class Stiffness {
   public:
    StiffnessMode mode;
    Assembly<nlayer> *ass{nullptr};
    Stiffness(std::vector<Particle<nlayer> *> &, StiffnessMode m) : mode(m) {}
    void reset(std::vector<Particle<nlayer> *> &) {}
    // localStiffness picks ANA or FD by mode (stiffness.h:50-56); FD is refused for J2Plasticity
    void calcStiffness(std::vector<Particle<nlayer> *> &) {
        lpm::check(lpm_set_mode(ass->dev->ctx, "stiffness_fd", mode == StiffnessMode::FiniteDifference ? 1 : 0),
                   ass->dev->ctx, "StiffnessMode");
        lpm::check(lpm_calc_stiffness(ass->dev->ctx), ass->dev->ctx, "calcStiffness");
    }
    void updateStiffnessDispBC(std::vector<Particle<nlayer> *> &) {
        lpm::check(lpm_update_stiffness_disp_bc(ass->dev->ctx), ass->dev->ctx, "updateStiffnessDispBC");
    }
    // the matrix handed to CG as scalar CSR (FP64 values)
    void toCSR(std::vector<int32_t> &rowptr, std::vector<int32_t> &col, std::vector<double> &val) {
        const int64_t nnz = lpm_size(ass->dev->ctx, "knnz");
        rowptr.resize(lpm_size(ass->dev->ctx, "ndof") + 1);
        col.resize(nnz);
        val.resize(nnz);
        lpm::check(lpm_download_matrix(ass->dev->ctx, rowptr.data(), col.data(), val.data(), nnz), ass->dev->ctx, "K");
    }
};

// ------------------------------------------------------------------ solver.h
template <int nlayer>
class Solver {
   public:
    int problem_size, max_NR_iter;
    double tol_NR_iter;
    std::string dumpFile;
    SolverMode sol_mode;
    std::vector<double> disp;
    std::vector<double> reaction_force;
    Stiffness<nlayer> stiffness;
    Assembly<nlayer> ass;

    Solver(Assembly<nlayer> &p_ass, const StiffnessMode &sm, const SolverMode &solm, const std::string &p_dump, const int &p_niter,
           const double &p_tol)
        : problem_size(p_ass.pt_sys[0]->cell.dim * (int)p_ass.pt_sys.size()),
          max_NR_iter(p_niter),
          tol_NR_iter(p_tol),
          dumpFile(p_dump),
          sol_mode(solm),
          disp(problem_size, 0.0),
          stiffness(p_ass.pt_sys, sm),
          ass(p_ass) {
        ass.ensureDevice();
        p_ass.dev = ass.dev;
        stiffness.ass = &ass;
        lpm_set_mode(ass.dev->ctx, "max_NR_iter", max_NR_iter);
        lpm_set_mode(ass.dev->ctx, "tol_NR", tol_NR_iter);
        lpm::check(lpm_set_mode(ass.dev->ctx, "stiffness_fd", sm == StiffnessMode::FiniteDifference ? 1 : 0), ass.dev->ctx,
                   "StiffnessMode");
    }

    lpm_ctx *ctx() { return ass.dev->ctx; }

    // LoadStep -> lpm_load_step (groups registered once on the device)
    struct CStep {
        std::vector<lpm_disp_bc> d;
        std::vector<lpm_force_bc> f;
        lpm_load_step view() const { return {(int32_t)d.size(), d.data(), (int32_t)f.size(), f.data()}; }
    };
    CStep convert(LoadStep<nlayer> &ls) {
        CStep c;
        for (auto &d : ls.dispBCs)
            c.d.push_back({ass.groupId(d.group), d.load_mode == LoadMode::Relative ? LPM_RELATIVE : LPM_ABSOLUTE, (int32_t)d.flag, d.step});
        for (auto &f : ls.forceBCs)
            c.f.push_back({ass.groupId(f.group), f.load_mode == LoadMode::Relative ? LPM_RELATIVE : LPM_ABSOLUTE, f.fx, f.fy, f.fz});
        return c;
    }

    void updateDisplacementBC(LoadStep<nlayer> &ls) {
        LoadStep<nlayer> only = ls;
        only.forceBCs.clear();
        applyBC(only);
    }
    void updateForceBC(LoadStep<nlayer> &ls) {
        LoadStep<nlayer> only = ls;
        only.dispBCs.clear();
        applyBC(only);
    }
    void applyBC(LoadStep<nlayer> &ls) {
        ass.push();
        CStep c = convert(ls);
        lpm_load_step v = c.view();
        lpm::check(lpm_apply_bc(ctx(), &v), ctx(), "apply BC");
        ass.pull(false);
    }
    // Solver::updateRR (solver.h:108-128)
    void updateRR() {
        double a = 0, b = 0;
        lpm::check(lpm_update_rr(ctx(), &a, &b), ctx(), "updateRR");
        int k = lpm_download(ctx(), "reaction", nullptr, 0);
        reaction_force.assign(k > 0 ? k : 0, 0.0);
        if (k > 0) lpm::check(lpm_download(ctx(), "reaction", reaction_force.data(), k), ctx(), "reaction");
    }
    void LPM_CG() { solveLinearSystem(); }
    void solveLinearSystem() {
        int32_t it = 0;
        double err = 0;
        lpm::check(lpm_solve_linear_system(ctx(), &it, &err), ctx(), "solveLinearSystem");
        lpm::check(lpm_download(ctx(), "disp", disp.data(), (int64_t)disp.size()), ctx(), "disp");
    }
    int NewtonIteration() {
        ass.push();
        int32_t ni = 0;
        lpm::check(lpm_newton_iteration(ctx(), &ni), ctx(), "NewtonIteration");
        ass.pull(false);
        return ni;
    }
};

// ------------------------------------------------------------------ solver_static.h
template <int nlayer>
class SolverStatic : public Solver<nlayer> {
   public:
    int undamaged_pt_type{0};
    SolverStatic(const int &p_undamaged, Assembly<nlayer> &p_ass, const StiffnessMode &sm, const SolverMode &solm,
                 const std::string &p_dump, const int &p_niter, const double &p_tol)
        : Solver<nlayer>{p_ass, sm, solm, p_dump, p_niter, p_tol}, undamaged_pt_type{p_undamaged} {}

    bool updateStaticDamage() {
        this->ass.push();
        int32_t any = 0;
        lpm::check(lpm_update_static_damage(this->ctx(), undamaged_pt_type, &any), this->ctx(), "updateStaticDamage");
        this->ass.pull();
        return any != 0;
    }

    // SolverStatic::solveProblemStep (solver_static.h:97-149)
    bool solveProblemStep(LoadStep<nlayer> &load_step, double &dt) {
        (void)dt;
        auto c = this->convert(load_step);
        lpm_load_step v = c.view();
        int32_t ok = 0;
        lpm::check(lpm_solve_step_static(this->ctx(), &v, undamaged_pt_type, &ok), this->ctx(), "solveProblemStep");
        return ok != 0;
    }

    // SolverStatic::solveProblem (solver_static.h:64-95)
    void solveProblem(std::vector<LoadStep<nlayer>> &load, int &start_index) {
        this->ass.push();
        int n_step = (int)load.size();
        double dt = 1;
        std::vector<double> pex;
        for (int i = 0; i < n_step; i++) {
            for (int halvings = 0;; ++halvings) {
                if (halvings > 30)  // the reference halves without end
                    throw std::runtime_error("load step " + std::to_string(i + start_index) + " does not converge after 30 halvings");
                printf("Loading step-%d, iteration starts:\n", i + start_index);
                double t1 = omp_get_wtime();
                if (lpm::options().rollback_pex) {
                    pex.resize(3 * this->ass.pt_sys.size());
                    lpm_download(this->ctx(), "Pex", pex.data(), (int64_t)pex.size());
                }
                if (solveProblemStep(load[i], dt)) {
                    if (lpm::options().async_dump) {
                        this->ass.writeDumpAsync(this->dumpFile, i + start_index);
                    } else {
                        this->ass.pull(false);
                        this->ass.writeDump(this->dumpFile, i + start_index);
                    }
                    printf("Loading step %d has finished, spent %f seconds\n\nData output ...\n\n", i + start_index,
                           omp_get_wtime() - t1);
                    break;
                }
                lpm::check(lpm_reset_state(this->ctx(), 1), this->ctx(), "resetStateVar");
                if (lpm::options().rollback_pex) lpm_upload(this->ctx(), "Pex", pex.data(), (int64_t)pex.size());
                load[i].loadCutHalf();
                load.insert(load.begin() + i, load[i]);
                ++n_step;
                printf("Step-%d not converging\n\n", i + start_index);
            }
        }
        this->ass.waitDump();
        this->ass.pull(true);
        start_index += n_step;
    }
};

// ------------------------------------------------------------------ solver_fatigue.h
inline int func_dNdt_linear(const double &tau) { return (int)(1 / tau); }
inline int func_dNdt_exponental(const double &tau, const double &t) { return (int)(1 / tau * std::exp(t / tau)); }

template <int nlayer>
class SolverFatigue : public Solver<nlayer> {
   public:
    int undamaged_pt_type{0};
    TimeMapMode t_mode;
    double tau{0};
    std::vector<double> load_spectrum;

    SolverFatigue(const int &p_undamaged, Assembly<nlayer> &p_ass, const StiffnessMode &sm, const SolverMode &solm,
                  const TimeMapMode &tm, const double &p_tau, const std::string &p_dump, const int &p_niter, const double &p_tol)
        : Solver<nlayer>{p_ass, sm, solm, p_dump, p_niter, p_tol}, undamaged_pt_type{p_undamaged}, t_mode(tm), tau(p_tau) {}

    // SolverFatigue::readLoad (solver_fatigue.h:40-63)
    void readLoad(const std::string &loadFile) {
        FILE *f = fopen(loadFile.c_str(), "r+");
        if (f == NULL) {
            printf("\'%s\' does not exist!\n", loadFile.c_str());
            exit(1);
        }
        double x;
        while (fscanf(f, "%lf", &x) > 0) load_spectrum.push_back(x);
        fclose(f);
    }

    bool updateFatigueDamage(double dNdt) {
        this->ass.push();
        int32_t any = 0;
        lpm::check(lpm_update_fatigue_damage(this->ctx(), dNdt, undamaged_pt_type, &any), this->ctx(), "updateFatigueDamage");
        this->ass.pull();
        return any != 0;
    }

    bool solveProblemStep(LoadStep<nlayer> &ls) {
        auto c = this->convert(ls);
        lpm_load_step v = c.view();
        int32_t ok = 0;
        lpm::check(lpm_solve_step_fatigue(this->ctx(), &v, &ok), this->ctx(), "solveProblemStep");
        return ok != 0;
    }

    // SolverFatigue::solveProblemStatic (solver_fatigue.h:243-275)
    void solveProblemStatic(std::vector<LoadStep<nlayer>> &load, int &start_index) {
        this->ass.push();
        int n_step = (int)load.size();
        for (int i = 0; i < n_step; i++) {
            for (int halvings = 0;; ++halvings) {
                if (halvings > 30)  // the reference halves without end
                    throw std::runtime_error("load step " + std::to_string(i + start_index) + " does not converge after 30 halvings");
                printf("Loading step-%d, iteration starts:\n", i + start_index);
                if (solveProblemStep(load[i])) break;
                lpm::check(lpm_reset_state(this->ctx(), 1), this->ctx(), "resetStateVar");
                load[i].loadCutHalf();
                load.insert(load.begin() + i, load[i]);
                ++n_step;
                printf("Step-%d not converging\n\n", i + start_index);
            }
            lpm::check(lpm_update_state_var(this->ctx()), this->ctx(), "updateStateVar");
            lpm::check(lpm_store_state(this->ctx()), this->ctx(), "storeStateVar");
            if (lpm::options().async_dump) {
                this->ass.writeDumpAsync(this->dumpFile, i);
            } else {
                this->ass.pull(false);
                this->ass.writeDump(this->dumpFile, i);
            }
        }
        this->ass.waitDump();
        this->ass.pull(true);
        start_index += n_step;
    }

    // SolverFatigue::generateLoadCycle (solver_fatigue.h:128-180)
    std::vector<LoadStep<nlayer>> generateLoadCycle(FatigueLoadType lt, int N, int n_interval,
                                                    std::vector<std::vector<Particle<nlayer> *>> g) {
        double f0 = load_spectrum[2 * N], f1 = load_spectrum[2 * N + 1], f2 = load_spectrum[2 * N + 2];
        std::vector<LoadStep<nlayer>> cyc;
        auto add = [&](double v) {
            if (lt == FatigueLoadType::LoadUniaxialDisp) cyc.push_back(LoadUniaxialDisp<nlayer>(v, g[0], g[1], g[2]));
            if (lt == FatigueLoadType::LoadUniaxialForce) cyc.push_back(LoadUniaxialForce<nlayer>(v, g[0], g[1], g[2]));
            if (lt == FatigueLoadType::LoadDogBoneDisp) cyc.push_back(LoadDogBoneDisp<nlayer>(0, v, 0, g[0], g[1]));
            if (lt == FatigueLoadType::LoadCTForce) cyc.push_back(LoadCTForce<nlayer>(v, g[0], g[1], g[2]));
        };
        for (int i = 0; i < n_interval; ++i) add((f1 - f0) / n_interval);
        for (int i = 0; i < n_interval; ++i) add((f2 - f1) / n_interval);
        return cyc;
    }

    // SolverFatigue::solveProblemOneCycle (solver_fatigue.h:220-241)
    void solveProblemOneCycle(std::vector<LoadStep<nlayer>> &load, double dNdt) {
        for (size_t i = 0; i < load.size(); ++i) {
            printf("Substep-%d, iteration starts:\n", (int)i);
            solveProblemStep(load[i]);
            lpm::check(lpm_update_state_var(this->ctx()), this->ctx(), "updateStateVar");
        }
        int32_t any = 0;
        lpm::check(lpm_update_fatigue_damage(this->ctx(), dNdt, undamaged_pt_type, &any), this->ctx(), "updateFatigueDamage");
        lpm::check(lpm_store_state(this->ctx()), this->ctx(), "storeStateVar");
        lpm::check(lpm_update_geometry(this->ctx()), this->ctx(), "updateGeometry");
        lpm::check(lpm_update_force_state(this->ctx()), this->ctx(), "updateForceState");
    }

    // SolverFatigue::solveProblemCyclic (solver_fatigue.h:182-218)
    void solveProblemCyclic(FatigueLoadType lt, int n_interval, std::vector<std::vector<Particle<nlayer> *>> g) {
        this->ass.push();
        double n_cycles = 0.5 * (load_spectrum.size() - 1);
        double t{0}, dNdt{1};
        int N{0};
        do {
            printf("Cycle-%d starts:\n\n", N + 1);
            if ((size_t)(2 * N + 2) >= load_spectrum.size()) throw std::runtime_error("load spectrum too short");
            std::vector<LoadStep<nlayer>> cyc = generateLoadCycle(lt, N, n_interval, g);
            dNdt = (t_mode == TimeMapMode::Linear) ? func_dNdt_linear(tau) : func_dNdt_exponental(tau, t++);
            solveProblemOneCycle(cyc, dNdt);
            this->ass.pull(false);
            double ave_u{0};
            int n_pt = 0;
            for (auto *pt : this->ass.pt_sys)
                if (pt->type != undamaged_pt_type) {
                    n_pt++;
                    ave_u += pt->state_var[0];
                }
            printf("Average strain energy is: %g\n", ave_u / n_pt);
            N += (int)dNdt;
            this->ass.writeDump(this->dumpFile, N);
        } while (N < n_cycles);
        this->ass.pull(true);
    }
};
