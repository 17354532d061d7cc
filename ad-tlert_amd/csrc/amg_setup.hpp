// Host-side symbolic setup of the smoothed-aggregation multigrid preconditioner.
//
// One hierarchy is built per CSR pattern from a representative matrix of that pattern
// (strength of connection and the damping weights need values).  Each level l holds the
// pattern of A_l, its aggregation, the pattern of the smoothed prolongator
//     P_l = (I - omega_l D_F^-1 A_F) P0_l
// (P0_l piecewise constant on aggregates; A_F the filtered matrix: strong off-diagonal
// entries only, the weak ones lumped into the diagonal D_F), its transpose (restriction), and
// the gather lists with which the device recomputes, for every system and every solve, the
// numeric P_l, T_l = A_l P_l and A_{l+1} = P_l^T T_l in a fixed order (so the Galerkin
// operators are bitwise repeatable).  Filtering keeps the coarse operators sparse in 3D.
//
// Method: P. Vanek, J. Mandel, M. Brezina, "Algebraic multigrid by smoothed aggregation
// for second and fourth order elliptic problems", Computing 56 (1996) 179-196 (aggregation,
// prolongator smoothing with the filtered matrix).
#pragma once

#include <cstdint>
#include <vector>

namespace adt {
namespace amg {

struct HostLevel {
    int n = 0;   // rows of A_l
    int nc = 0;  // rows of A_{l+1} (aggregates); 0 on the coarsest level
    std::vector<int> ptr, idx, diag;  // pattern of A_l and the slot of each diagonal entry
    std::vector<double> rep;          // representative values of A_l
    double omega = 2.0 / 3.0;         // smoother damping s * 4 / (3 lambda_max(D^-1 A)), s = smoother_scale
    double omega_p = 2.0 / 3.0;       // prolongation damping 4 / (3 lambda_max(D_F^-1 A_F))
    std::vector<int> agg;             // aggregate of each row, -1: left out (no strong connection)
    std::vector<double> pfix;         // a given prolongator's values (empty: smoothed aggregation)
    std::vector<int> dmap_ptr, dmap;  // per row: slots summed into D_F (diagonal + weak entries)
    // prolongator P (n x nc)
    std::vector<int> pptr, pidx;
    std::vector<int> pmap_ptr, pmap;  // per P slot: strong off-diagonal A slots of row i in aggregate J
    std::vector<int> pself;           // per P slot: 1 if J is the row's own aggregate
    // restriction R = P^T (nc x n): per coarse row, (fine row, P slot) entries in fine-row order
    std::vector<int> rptr, rrow, rslot;
    // T = A P (n x nc): pattern and (A slot, P slot) product pairs per T slot
    std::vector<int> tptr, tidx, tpair_ptr, tpair_a, tpair_p;
    // A_{l+1} = R T: (P slot, T slot) product pairs per slot of the next level's pattern
    std::vector<int> cpair_ptr, cpair_p, cpair_t;
};

// A given prolongator for the finest level (CSR over the fine rows, nc coarse columns), e.g.
// the interpolation of ten-node (P2) tetrahedra from their vertices (P1): vertex rows copy their
// vertex, edge-midpoint rows average the edge's two ends.  Coarser levels aggregate as usual.
struct FixedProlongator {
    int nc = 0;
    const int32_t* ptr = nullptr;
    const int32_t* idx = nullptr;
    const double* val = nullptr;
};

struct Options {
    double theta = 0.08;    // strength threshold |a_ij| >= theta sqrt(a_ii a_jj)
    int coarse_max = 256;   // stop coarsening at this many rows (dense inverse on the coarsest level)
    int max_levels = 12;
    // Jacobi smoother damping of the finest level in units of 4 / (3 lambda_max(D^-1 A)), the
    // choice that minimizes the largest amplification of the upper half of the spectrum for a
    // stand-alone smoother.  Inside a Krylov-accelerated cycle a larger damping is better; the
    // cycle stays positive definite while omega lambda_max < 2 on every system of the pattern,
    // i.e. smoother_scale < 1.5 / (lambda_max of a system / lambda_max of the representative).
    // The finest operators of a pattern vary little from system to system (within 1% over
    // twelve conductivity models of the P2 terrain mesh; below the representative's over the
    // wavenumbers of a 2.5D pattern); the coarse Galerkin operators vary more (35% on the first
    // coarse level of the terrain mesh), so coarse levels keep the unscaled damping.
    double smoother_scale = 1.0;
    const FixedProlongator* first = nullptr;  // finest-level prolongator (null: aggregation)
};

// Builds the levels from a CSR pattern and representative values (rows = n).  The last
// level is the coarsest (dense inverse).  Returns an empty vector if the input is invalid.
std::vector<HostLevel> build_hierarchy(int n, const int32_t* ptr, const int32_t* idx, const double* values,
                                       const Options& opt);

}  // namespace amg
}  // namespace adt
