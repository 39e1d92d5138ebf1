// Host-side setup of a particle system: lattice points, unit-cell constants and the
// directed bond list. Part of liblpm_mi355x.so (C ABI in include/lpm_mi355x.h).
//
// The bond search replaces Assembly::createBonds (include/assembly.h:212-234), which tests
// all N^2 pairs. Here particles are binned on a uniform grid whose cell edge is the bond
// cutoff, and only the 27 (3D) or 9 (2D) surrounding cells are examined. Candidates are
// visited in increasing id and accepted with the reference's own test, evaluated with the
// reference's own expression, so the lists, their order and the bond lengths are identical
// to the all-pairs loop.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "lpm_mi355x.h"

namespace {

struct CellConst {
    int dim, nneighbors, nn1, nn2, nafem;
    double cutoff[2], volume;
    double data[9];  // el_mapping source array, in the order it is written in unit_cell.h
};

// UnitCell::setSquare/setHexagon/setSimpleCubic/setFCC/setBCC (unit_cell.h:44-127)
bool unit_cell(int lattice, double r, CellConst &c) {
    std::memset(&c, 0, sizeof c);
    if (lattice == LPM_SQUARE2D) {
        const double d[9] = {1. / 2., -1. / 2., 0., 0., 0., 1. / 2., 0., 1. / 12., -1. / 12.};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 2; c.nneighbors = 8; c.nn1 = 4; c.nn2 = 4; c.nafem = 16;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 2.0 * std::sqrt(2.0) * r;
        c.volume = 4 * std::pow(r, 2);
    } else if (lattice == LPM_HEXAGON2D) {
        const double d[9] = {std::sqrt(3.0) / 12.0, -std::sqrt(3.0) / 12.0, 0., -std::sqrt(3.0) / 144.0,
                             std::sqrt(3.0) / 48.0, 0.0, 0.0, 0.0, 1.0};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 2; c.nneighbors = 12; c.nn1 = 6; c.nn2 = 6; c.nafem = 30;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 2.0 * std::sqrt(3.0) * r;
        c.volume = 2 * std::sqrt(3) * std::pow(r, 2);
    } else if (lattice == LPM_SC3D) {
        const double d[9] = {1, -1, -1.0, 0.0, 0.0, 1., 0.0, 1 / 18.0, -1 / 18.0};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 3; c.nneighbors = 18; c.nn1 = 6; c.nn2 = 12; c.nafem = 60;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 2.0 * std::sqrt(2.0) * r;
        c.volume = std::pow(2 * r, 3);
    } else if (lattice == LPM_FCC3D) {
        const double d[9] = {0.0, 0.0, std::sqrt(2.0), std::sqrt(2.0) / 4, -std::sqrt(2.0) / 4, -std::sqrt(2.0) / 4,
                             0.0, std::sqrt(2.0) / 24, -std::sqrt(2.0) / 24};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 3; c.nneighbors = 18; c.nn1 = 12; c.nn2 = 6; c.nafem = 60;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 2.0 * std::sqrt(2.0) * r;
        c.volume = 4.0 * std::sqrt(2.0) * std::pow(r, 3);
    } else if (lattice == LPM_BCC3D) {
        const double d[9] = {0.0, 0.0, std::sqrt(3.0), 1 / std::sqrt(3.0), -1 / std::sqrt(3.0), 0.,
                             0.0, std::sqrt(3.0) / 14, std::sqrt(3.0) / 14};
        std::memcpy(c.data, d, sizeof d);
        c.dim = 3; c.nneighbors = 14; c.nn1 = 8; c.nn2 = 6; c.nafem = 40;
        c.cutoff[0] = 2.0 * r; c.cutoff[1] = 4.0 / std::sqrt(3.0) * r;
        c.volume = 32.0 * std::sqrt(3.0) / 9.0 * std::pow(r, 3);
    } else {
        return false;
    }
    return true;
}

// createRMatrix (utilities.h:8-82): R[3*row + col]
void rotation(int flag, const double *a, double *R) {
    const double a1 = a[0], a2 = a[1], a3 = a[2];
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

// R * p as Eigen's column-major GEMV accumulates it (column 0, then 1, then 2)
inline void rotate(const double *R, const double *p, double *o) {
    for (int r = 0; r < 3; ++r) {
        double s = R[3 * r] * p[0];
        s = s + R[3 * r + 1] * p[1];
        s = s + R[3 * r + 2] * p[2];
        o[r] = s;
    }
}

struct Sink {
    double *out;
    int64_t cap, n = 0;
    void add(const double *q) {
        if (out && n < cap) std::memcpy(out + 3 * n, q, 3 * sizeof(double));
        ++n;
    }
};

inline bool inside(const double *q, const double *box, bool three) {
    bool ok = q[0] >= box[0] && q[0] <= box[1] && q[1] >= box[2] && q[1] <= box[3];
    return three ? ok && q[2] >= box[4] && q[2] <= box[5] : ok;
}

}  // namespace

extern "C" {

int lpm_unit_cell(int lattice, double radius, int literal, double *out, double *map) {
    CellConst c;
    if (!unit_cell(lattice, radius, c) || !out) return LPM_E_ARG;
    out[0] = c.dim; out[1] = c.nneighbors; out[2] = c.nn1; out[3] = c.nn2; out[4] = c.nafem;
    out[5] = c.cutoff[0]; out[6] = c.cutoff[1]; out[7] = c.volume;
    if (map)
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) map[3 * r + q] = literal ? c.data[3 * q + r] : c.data[3 * r + q];
    return LPM_OK;
}

int lpm_rotation_matrix(int eulerflag, const double *angles, double *R) {
    if (!angles || !R || eulerflag < 0 || eulerflag > 2) return LPM_E_ARG;
    rotation(eulerflag, angles, R);
    return LPM_OK;
}

int lpm_generate_lattice(int lattice, const double *box, double radius, int eulerflag, const double *angles, double *xyz,
                         int64_t cap, int64_t *n) {
    if (!box || !angles || !n || radius <= 0) return LPM_E_ARG;
    double R[9];
    rotation(eulerflag, angles, R);
    return lpm_generate_lattice_r(lattice, box, radius, R, xyz, cap, n);
}

// createPlateSQ2D / createPlateHEX2D / createCuboidSC3D / createCuboidFCC3D (utilities.h:84-301)
int lpm_generate_lattice_r(int lattice, const double *box, double radius, const double *R, double *xyz, int64_t cap,
                           int64_t *n) {
    if (!box || !R || !n || radius <= 0) return LPM_E_ARG;
    Sink s{xyz, xyz ? cap : 0};
    double p[3] = {0, 0, 0}, q[3];
    if (lattice == LPM_SQUARE2D || lattice == LPM_HEXAGON2D) {
        const double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2));
        const double bt[4] = {-a, a, -a, a};
        if (lattice == LPM_SQUARE2D) {
            const double hx = 2 * radius, hy = hx;
            const int pfr = (int)std::floor((bt[1] - bt[0]) / hx), rows = (int)std::floor((bt[3] - bt[2]) / hy);
            for (int j = 1; j <= rows; j++) {
                p[1] = bt[2] + hy * (j - 1) + radius;
                for (int i = 1; i <= pfr; i++) {
                    p[0] = bt[0] + hx * (i - 1) + radius;
                    rotate(R, p, q);
                    if (inside(q, box, false)) { const double v[3] = {q[0], q[1], 0}; s.add(v); }
                }
            }
        } else {
            const double hx = 2 * radius, hy = hx * std::sqrt(3.0) / 2.0;
            const int pfr = 1 + (int)std::floor((bt[1] - bt[0]) / hx), rows = 1 + (int)std::floor((bt[3] - bt[2]) / hy);
            for (int j = 1; j <= rows; j++) {
                p[1] = bt[2] + hy * (j - 1);
                const int cnt = (j % 2 == 1) ? pfr : pfr - 1;
                for (int i = 1; i <= cnt; i++) {
                    p[0] = (j % 2 == 1) ? bt[0] + hx * (i - 1) : bt[0] + hx * (2 * i - 1) / 2.0;
                    p[2] = 0;
                    rotate(R, p, q);
                    if (inside(q, box, false)) s.add(q);
                }
            }
        }
    } else if (lattice == LPM_SC3D || lattice == LPM_FCC3D) {
        const double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2) + std::pow(box[5] - box[4], 2));
        const double bt[6] = {-a, a, -a, a, -a, a};
        if (lattice == LPM_SC3D) {
            const double hx = 2 * radius, hy = hx, hz = hx;
            const int pfr = 1 + (int)std::floor((bt[1] - bt[0]) / hx), rows = 1 + (int)std::floor((bt[3] - bt[2]) / hy),
                      layers = 1 + (int)std::floor((bt[5] - bt[4]) / hz);
            for (int k = 1; k <= layers; k++) {
                p[2] = bt[4] + hz * (k - 1);
                for (int j = 1; j <= rows; j++) {
                    p[1] = bt[2] + hy * (j - 1);
                    for (int i = 1; i <= pfr; i++) {
                        p[0] = box[0] + hx * (i - 1);  // x starts at box[0] in the reference (utilities.h:195)
                        rotate(R, p, q);
                        if (inside(q, box, true)) s.add(q);
                    }
                }
            }
        } else {
            const double hx = 4 * radius / std::sqrt(2.0), hy = hx, hz = hx;
            const int pfr = 1 + (int)std::floor((bt[1] - bt[0]) / hx), rows = 1 + (int)std::floor((bt[3] - bt[2]) / (hy / 2.0)),
                      layers = 1 + (int)std::floor((bt[5] - bt[4]) / (hz / 2.0));
            for (int k = 1; k <= layers; k++) {
                p[2] = bt[4] + hz / 2.0 * (k - 1);
                for (int j = 1; j <= rows; j++) {
                    p[1] = bt[2] + hy / 2.0 * (j - 1);
                    const bool full = ((k % 2 == 1) == (j % 2 == 1));
                    const int cnt = full ? pfr : pfr - 1;
                    for (int i = 1; i <= cnt; i++) {
                        p[0] = full ? bt[0] + hx * (i - 1) : bt[0] + hx * (i - 1) + hx / 2.0;
                        rotate(R, p, q);
                        if (inside(q, box, true)) s.add(q);
                    }
                }
            }
        }
    } else {
        return LPM_E_ARG;
    }
    *n = s.n;
    if (xyz && s.n > cap) return LPM_E_LIMIT;
    return LPM_OK;
}

int lpm_build_bonds(int lattice, double radius, int64_t n, const double *xyz, int32_t *bond_ptr, int32_t *bond_split,
                    int32_t *bond_p2, int64_t cap, int64_t *nbonds) {
    CellConst cc;
    if (!unit_cell(lattice, radius, cc) || n < 0 || (n > 0 && !xyz) || !nbonds) return LPM_E_ARG;
    if (n > 0x7ffffffeLL) return LPM_E_LIMIT;
    const double c1 = 1.01 * cc.cutoff[1], c0 = 1.01 * cc.cutoff[0];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], xyz[3 * i + k]);
            hi[k] = std::max(hi[k], xyz[3 * i + k]);
        }
    const double h = c1 * 1.0001;
    int64_t nc[3];
    for (int k = 0; k < 3; ++k) nc[k] = n > 0 ? std::max<int64_t>(1, (int64_t)std::floor((hi[k] - lo[k]) / h) + 1) : 1;
    const int64_t ncell = nc[0] * nc[1] * nc[2];
    auto cell_of = [&](const double *x, int64_t *c) {
        for (int k = 0; k < 3; ++k) {
            c[k] = (int64_t)std::floor((x[k] - lo[k]) / h);
            c[k] = std::min(std::max<int64_t>(c[k], 0), nc[k] - 1);
        }
    };
    std::vector<int64_t> head(ncell + 1, 0), pcell(n);
    std::vector<int32_t> items(n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3];
        cell_of(&xyz[3 * i], c);
        pcell[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
        head[pcell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) head[c + 1] += head[c];
    {
        std::vector<int64_t> fill(head.begin(), head.end() - 1);
        for (int64_t i = 0; i < n; ++i) items[fill[pcell[i]]++] = (int32_t)i;  // ids ascend within a cell
    }
    std::vector<std::vector<int32_t>> l0(n), l1(n);
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3];
        cell_of(&xyz[3 * i], c);
        std::vector<int32_t> cand;
        cand.reserve(64);
        for (int64_t dz = -1; dz <= 1; ++dz)
            for (int64_t dy = -1; dy <= 1; ++dy)
                for (int64_t dx = -1; dx <= 1; ++dx) {
                    const int64_t x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= nc[0] || y >= nc[1] || z >= nc[2]) continue;
                    const int64_t cell = (z * nc[1] + y) * nc[0] + x;
                    for (int64_t t = head[cell]; t < head[cell + 1]; ++t) cand.push_back(items[t]);
                }
        std::sort(cand.begin(), cand.end());
        const double *a = &xyz[3 * i];
        for (int32_t j : cand) {
            const double *b = &xyz[3 * (int64_t)j];
            // Particle::distanceTo (particle.h:276-280)
            const double distance = std::sqrt(std::pow((a[0] - b[0]), 2) + std::pow((a[1] - b[1]), 2) + std::pow((a[2] - b[2]), 2));
            if ((distance < c1) && (j != i)) (distance < c0 ? l0[i] : l1[i]).push_back(j);
        }
    }
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) total += (int64_t)(l0[i].size() + l1[i].size());
    *nbonds = total;
    if (total > 0x7fffffffLL) return LPM_E_LIMIT;
    if (!bond_p2) return LPM_OK;
    if (!bond_ptr || !bond_split || cap < total) return LPM_E_ARG;
    int64_t o = 0;
    for (int64_t i = 0; i < n; ++i) {
        bond_ptr[i] = (int32_t)o;
        for (int32_t j : l0[i]) bond_p2[o++] = j;
        bond_split[i] = (int32_t)o;
        for (int32_t j : l1[i]) bond_p2[o++] = j;
    }
    bond_ptr[n] = (int32_t)o;
    return LPM_OK;
}

// benchmark/AM_Ti64/generate_dump.ipynb, createPlate: a particle at the centre of every pixel
// equal to 255, rows from the top of the image down, in row-major order
int lpm_image_particles(const uint8_t *img, int64_t nrows, int64_t ncols, double pixel, double *xyz, int64_t cap,
                        int64_t *n) {
    if (!img || nrows < 0 || ncols < 0 || !(pixel > 0) || !n) return LPM_E_ARG;
    const double radius = pixel / 2;
    const double size_y = nrows * pixel;
    int64_t m = 0;
    for (int64_t j = 0; j < nrows; ++j) {
        const double y = size_y - (radius + pixel * j);
        for (int64_t i = 0; i < ncols; ++i) {
            if (img[j * ncols + i] != 255) continue;
            if (xyz) {
                if (m >= cap) return LPM_E_LIMIT;
                xyz[3 * m] = radius + pixel * i;
                xyz[3 * m + 1] = y;
                xyz[3 * m + 2] = 0.0;
            }
            ++m;
        }
    }
    *n = m;
    return LPM_OK;
}

int lpm_ce_from_e(int plane_stress, double E, double mu, double *Ce) {
    if (!Ce) return LPM_E_ARG;
    // particle_elastic.h:77-88
    if (!plane_stress) {
        Ce[0] = E * (1.0 - mu) / (1.0 + mu) / (1.0 - 2.0 * mu);
        Ce[1] = E * mu / (1.0 + mu) / (1.0 - 2.0 * mu);
        Ce[2] = E / 2.0 / (1.0 + mu);
    } else {
        Ce[0] = E / (1.0 - mu) / (1.0 + mu);
        Ce[1] = E * mu / (1.0 + mu) / (1.0 - mu);
        Ce[2] = E / 2.0 / (1.0 + mu);
    }
    return LPM_OK;
}

}  // extern "C"
