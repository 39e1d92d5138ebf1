// A driver written against the reference API (cf. reference examples/ex1_isoelasticity_2d_hex.cpp):
// 2D hexagonal plate, bottom edge clamped, top edge pulled in y over several load steps.
// Build: see examples/Makefile. Run: ./ex1_hex_plate [nx ny steps [tol [precision]]]
//   (precision 1: FP64 matrix and CG, relative CG tolerance 1e-13)
#include "lpm.h"
#include "particle.h"
#include "assembly.h"
#include "utilities.h"
#include "stiffness.h"
#include "load_step.h"
#include "solver_static.h"

#include <cstdlib>

int main(int argc, char **argv) {
    double start = omp_get_wtime();
    const int n_layer = 2;
    double radius = 0.5;
    int nx = argc > 1 ? atoi(argv[1]) : 20, ny = argc > 2 ? atoi(argv[2]) : 20, n_steps = argc > 3 ? atoi(argv[3]) : 3;
    UnitCell cell(LatticeType::Hexagon2D, radius);
    // Read el_mapping row-major (its documented meaning). The reference's column-major read
    // makes this lattice's tangent stiffness indefinite after the first step and CG diverges.
    lpm::options().elmap_literal = 0;
    if (argc > 5 && atoi(argv[5]) == 1) {
        lpm::options().precision = 1;
        lpm::options().cg_tol = 1e-13;
    }

    int eulerflag = 0;
    double angles[] = {PI / 180.0 * 0.0, PI / 180.0 * 0.0, PI / 180.0 * 0.0};
    Eigen::MatrixXd R_matrix = createRMatrix(eulerflag, angles);

    std::array<double, 2 * NDIM> box{0.0, 2 * radius * nx, 0.0, 2 * radius * 0.8660254037844386 * ny, 0.0, 1.0};
    std::vector<std::array<double, NDIM>> hex_xyz = createPlateHEX2D(box, cell, R_matrix);
    Assembly<n_layer> pt_ass{hex_xyz, box, cell, ParticleType::Elastic};
    printf("\nParticle number is %d\n", pt_ass.nparticle);

    bool is_plane_stress = false;
    double E0 = 69e3, mu0 = 0.3, critical_bstrain = 1;
    int max_iter = 30, start_index = 0;
    double tol_iter = argc > 4 ? atof(argv[4]) : 1e-10;
    int undamaged_pt_type = -1;
    std::string dumpFile{"ex1_hex_plate.dump"};

    double ymin = 1e300, ymax = -1e300;
    for (auto *p : pt_ass.pt_sys) { ymin = std::min(ymin, p->xyz[1]); ymax = std::max(ymax, p->xyz[1]); }
    std::vector<Particle<n_layer> *> top_group, bottom_group;
    for (Particle<n_layer> *p1 : pt_ass.pt_sys) {
        if (p1->xyz[1] > ymax - 1e-6) { top_group.push_back(p1); p1->type = 1; }
        if (p1->xyz[1] < ymin + 1e-6) { bottom_group.push_back(p1); p1->type = 2; }
        ParticleElastic<n_layer> *elpt = dynamic_cast<ParticleElastic<n_layer> *>(p1);
        elpt->setParticleProperty(0.0, is_plane_stress, E0, mu0, critical_bstrain);
    }

    std::vector<LoadStep<n_layer>> load;
    for (int i = 0; i < n_steps; i++) {
        LoadStep<n_layer> step;
        step.dispBCs.push_back(DispBC<n_layer>(bottom_group, LoadMode::Relative, 'x', 0.0));
        step.dispBCs.push_back(DispBC<n_layer>(bottom_group, LoadMode::Relative, 'y', 0.0));
        step.dispBCs.push_back(DispBC<n_layer>(top_group, LoadMode::Relative, 'y', 0.001));
        load.push_back(step);
    }

    pt_ass.updateGeometry();
    pt_ass.updateForceState();

    SolverStatic<n_layer> solv{undamaged_pt_type, pt_ass, StiffnessMode::Analytical, SolverMode::CG, dumpFile, max_iter, tol_iter};
    printf("Initialization finished in %f seconds\n\n", omp_get_wtime() - start);
    solv.solveProblem(load, start_index);

    // final top displacement and reaction (what a user of the reference reads back)
    double uy = 0, ry = 0;
    for (auto *p : top_group) { uy += p->xyz[1] - p->xyz_initial[1]; ry += p->Pin[1]; }
    printf("RESULT n=%d steps=%d top_uy=%.12e top_reaction_y=%.12e\n", pt_ass.nparticle, start_index, uy / top_group.size(), ry);
    printf("Computation time for total steps: %f seconds\n\n", omp_get_wtime() - start);
    return 0;
}
