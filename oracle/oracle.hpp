// CPU oracle: a plain restatement of the BoSSS algorithms on the DG-operator /
// block-solver hot path.  TEST INFRASTRUCTURE ONLY: the product library never
// links or calls this code; only tests/, the smoke check and the CPU-baseline
// leg of the benchmark load it, and only as the checker / CPU baseline.
//
// Every routine cites the reference file:line it follows (paths relative to the
// BoSSS source tree `src/`).  Abbreviations: F/ = L2-foundation/BoSSS.Foundation,
// P/ = L1-platform/BoSSS.Platform, AS/ = L3-solution/BoSSS.Solution.AdvancedSolvers.
#pragma once
#include <cstdint>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- basis ----
// Orthonormal full-polynomial basis on [-1,1]^D, degree-graded, within a degree
// ordered by ascending x-exponent, then y-exponent (F/Square.cs:192-260,
// F/Cube_polynomials.cs:13-60); psi = prod_d Ptilde_{e_d}(xi_d), Ptilde_n =
// sqrt((2n+1)/2) P_n (Legendre).
int basis_count(int D, int k);
void basis_exponents(int D, int k, std::vector<int>& ex);  // Np*D
void basis_eval(int D, int k, const double* xi, double* val, double* grad);

// Gauss-Legendre rule with n nodes on [-1,1] (F/LineQuadRules/Line{n}.txt).
void gauss_rule(int n, double* x, double* w);

// ----------------------------------------------------------------- grid ----
struct Grid {
    int D = 0;
    int n[3] = {1, 1, 1};
    std::vector<double> nodes[3];
    int J = 0;
    std::vector<int> ijk;              // J*3 logical index of each cell
    std::vector<double> lo, h;         // J*D box origin and extents
    // edges: interior (out >= 0) and boundary (out = -1)
    int E = 0;
    std::vector<int> e_in, e_out, e_dir, e_side;  // e_side: boundary side 2d+{0,1}, -1 interior
    std::vector<double> cj;            // F/GridData_CellData.cs:666-700
    int cell_index(int i, int j, int k) const;
};
Grid* grid_create(int D, const double* x, int nx1, const double* y, int ny1, const double* z, int nz1);

// ------------------------------------------------------------------ SIP ----
struct Sip {
    const Grid* g = nullptr;
    int D = 0, k = 0, Np = 0;
    int nq1 = 0;                  // Gauss points per direction
    double penalty_base = 1.3, penalty_deg = 0;
    std::vector<double> nu;       // per cell
    int bc[6] = {0, 0, 0, 0, 0, 0};  // 0 Dirichlet, 1 Neumann per boundary side
    std::vector<double> gbnd;     // per boundary edge node (layout: boundary edges in edge order x nodes)
    std::vector<int> bnd_edge_slot;  // E: first slot in gbnd, or -1
    int nqf = 0, nqv = 0;
    std::vector<double> f_rhs;    // N, RHS source field coefficients (optional)
    int chunk = 20;               // F/Quadrature.cs:43 CHUNK_LIMIT
    // cached reference basis values (volume nodes; face f = 2d+s nodes)
    std::vector<double> vpsi, vgpsi, fpsi, fgpsi;
};

// ------------------------------------------------------------------ BSR ----
struct Bsr {
    int nrows = 0, bs = 0;
    std::vector<int64_t> rowptr;
    std::vector<int> col;
    std::vector<double> val;      // nnzb*bs*bs, each block row-major
};

}  // namespace orc
