// A file-driven fatigue driver written against the reference API (cf. reference
// examples/ex5_Ti64_2d_fatigue_gage.cpp, which reads its spectrum with readLoad): the particle
// system comes from a dump file and a bond file (Assembly(dumpFile, bondFile, ...), readDump /
// readBond), an HCF plate is loaded statically to the spectrum's first value, then cycled through
// the spectrum file with cycle jumping.
// Build: see examples/Makefile.
// Run:   ./ex5_fatigue_files <prefix> <tau> <time_map 0|1> [A]
//   reads  <prefix>.dump, <prefix>.bond, <prefix>.load
//   writes <prefix>_out.bond (writeBond of what was read), <prefix>_search.bond (writeBond of
//          Assembly(dumpFile, ...), i.e. readDump + createBonds), <prefix>.final (binary FP64:
//          x, y, z, Pex x, y, z, damage per particle) and <prefix>_run.dump (the solver's frames)
#include "lpm.h"
#include "particle.h"
#include "assembly.h"
#include "utilities.h"
#include "stiffness.h"
#include "load_step.h"
#include "solver_fatigue.h"
#include "load_step_fatigue.h"

#include <cstdlib>

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s prefix tau time_map [A]\n", argv[0]);
        return 2;
    }
    const std::string prefix = argv[1];
    const double tau = atof(argv[2]);
    const TimeMapMode tmode = atoi(argv[3]) ? TimeMapMode::Exponential : TimeMapMode::Linear;
    const double f_A = argc > 4 ? atof(argv[4]) : 5.206e-5;

    const int n_layer = 2;
    const double radius = 0.5;
    UnitCell cell(LatticeType::Hexagon2D, radius);
    lpm::options().elmap_literal = 0;
    lpm::options().precision = 1;
    lpm::options().cg_tol = 1e-13;
    lpm::options().verbose = 0;
    lpm::options().async_dump = 0;

    Assembly<n_layer> pt_ass{prefix + ".dump", prefix + ".bond", cell, ParticleType::FatigueHCF};
    pt_ass.writeBond(prefix + "_out.bond");
    {
        Assembly<n_layer> searched{prefix + ".dump", cell, ParticleType::FatigueHCF};
        searched.writeBond(prefix + "_search.bond");
    }

    double xmin = 1e300, ymin = 1e300, ymax = -1e300;
    for (auto *p : pt_ass.pt_sys) {
        xmin = std::min(xmin, p->xyz[0]);
        ymin = std::min(ymin, p->xyz[1]);
        ymax = std::max(ymax, p->xyz[1]);
    }
    const int undamaged_pt_type = 4;
    std::vector<Particle<n_layer> *> top_group, left_group, bottom_group;
    for (Particle<n_layer> *p1 : pt_ass.pt_sys) {
        if (p1->xyz[0] < xmin + radius) left_group.push_back(p1), p1->type = 1;
        if (p1->xyz[1] > ymax - radius) top_group.push_back(p1), p1->type = undamaged_pt_type;
        if (p1->xyz[1] < ymin + radius) bottom_group.push_back(p1), p1->type = undamaged_pt_type;
        auto *ft = dynamic_cast<ParticleFatigueHCF<n_layer> *>(p1);
        ft->setParticleProperty(3.0, true, 110e3, 0.34, f_A, 2.019, 0.5, 1.0, 1.108);
    }
    pt_ass.searchNonlocalNeighbors(1.5);
    pt_ass.updateGeometry();
    pt_ass.updateForceState();

    SolverFatigue<n_layer> solv{undamaged_pt_type, pt_ass, StiffnessMode::Analytical, SolverMode::CG, tmode, tau,
                                prefix + "_run.dump", 30, 1e-9};
    solv.readLoad(prefix + ".load");

    std::vector<LoadStep<n_layer>> load_static;
    LoadStep<n_layer> step0;
    step0.dispBCs.push_back(DispBC<n_layer>(left_group, LoadMode::Relative, 'x', 0.0));
    step0.forceBCs.push_back(ForceBC<n_layer>(top_group, LoadMode::Relative, 0.0, solv.load_spectrum[0], 0.0));
    step0.dispBCs.push_back(DispBC<n_layer>(bottom_group, LoadMode::Relative, 'y', 0.0));
    load_static.push_back(step0);
    int start_index = 0;
    solv.solveProblemStatic(load_static, start_index);
    solv.solveProblemCyclic(FatigueLoadType::LoadUniaxialForce, 1, {left_group, top_group, bottom_group});

    FILE *f = fopen((prefix + ".final").c_str(), "wb");
    if (!f) return 1;
    for (auto *p : pt_ass.pt_sys) {
        double row[7] = {p->xyz[0], p->xyz[1], p->xyz[2], p->Pex[0], p->Pex[1], p->Pex[2], p->damage};
        fwrite(row, sizeof(double), 7, f);
    }
    fclose(f);
    printf("RESULT n=%d spectrum=%d\n", pt_ass.nparticle, (int)solv.load_spectrum.size());
    return 0;
}
