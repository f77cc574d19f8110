// Host-side lowering of a PI model to the fused static Liouvillian plan and
// to the device stencil tables.
//
// The fused plan has the reference's FusedStaticPIKernel content
// (src/liouvillian.jl:539-545): fused Hamiltonian and loss blocks per sector,
// collective gains, and one-body gain branches with their rectangular
// contractions.  The stencil tables re-express that plan, per output sector,
// as a list of two-sided sparse products  Y_l += s * A X_n B^H  that the
// device kernel evaluates one output coordinate per lane.
#pragma once

#include <complex>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace pimi {

using cd = std::complex<double>;

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

struct CSC {
    int64_t m = 0, n = 0;
    std::vector<int64_t> colptr;
    std::vector<int32_t> rows;
    std::vector<cd> vals;
    void zero(int64_t mm, int64_t nn) {
        m = mm; n = nn; colptr.assign(nn + 1, 0); rows.clear(); vals.clear();
    }
};

struct Branch {
    int32_t out = 0, in = 0;
    double r = 0.0;
    int32_t use_support = 0;
    std::vector<int32_t> srows, scols;
    std::vector<cd> svals;
    std::vector<cd> dense;  // m_out x m_in column-major (always kept)
};

using Partition = std::vector<int>;
using Pattern = std::vector<int>;  // GT entries row by row, row 1 first

// (parent index, zero-based local label, CG value) per child pattern
using Transitions = std::vector<std::vector<std::tuple<int32_t, int32_t, double>>>;

struct Geometry {
    int N = 0, d = 0;
    std::vector<Partition> sectors;
    std::vector<std::vector<Pattern>> patterns;
    std::vector<std::vector<Partition>> removed;  // children per sector (removable-corner order)
    std::map<std::pair<int, Partition>, Transitions> trans;
    std::vector<double> sqrt_f;                    // sqrt(RN(f^nu)) per sector

    void build(int N, int d);
    int64_t dim(int s) const { return (int64_t)patterns[s].size(); }
    std::vector<Partition> common(int li, int ni) const;
    double scale(int li, const Partition &mu, int ni) const;
    // tmp[a + c*m_l] = contraction of cell (a,c) for child mu with local operator X
    // (d x d column-major), summed in the packed-table order.
    void contract(int li, const Partition &mu, int ni, const cd *X, std::vector<cd> &tmp) const;
    CSC collective_block(int s, const cd *X) const;
};

// One jump channel of the model, in term order (src/trajectories.jl): its gain
// (kind 0: collective gain K[index]; 1: one-body gain ob[index]), rate and
// unscaled loss blocks L^dagger L per sector.
struct JumpChannel {
    int32_t kind = 0, index = 0;
    double rate = 0.0;
    std::vector<CSC> q;
};

struct FusedPlan {
    Geometry geom;
    std::vector<JumpChannel> channels;
    int ns = 0;
    int64_t n = 0;
    std::vector<int64_t> dims, offs;
    bool has_h = false, has_q = false;
    std::vector<CSC> H, Q;
    std::vector<cd> cg_scale;
    std::vector<std::vector<CSC>> K;
    std::vector<cd> ob_scale;
    std::vector<std::vector<Branch>> ob;
};

struct TermDesc {
    int kind;
    std::vector<cd> op, op2;  // d*d column-major
    double rate, hbar;
};

void lower_model(int N, int d, const std::vector<TermDesc> &terms, FusedPlan &out);

// Collective operator sum_i X_i as a PI coefficient vector.
void collective_operator(const FusedPlan &plan, const cd *X, cd *out);

// PI coefficients of A^{otimes N} (A: d x d column-major) by the one-box
// Clebsch-Gordan recurrence of iid_state.
void iid_state(const FusedPlan &plan, const cd *A, cd *out);

// Device stencil: per output sector, terms Y_l += s * A X_n B^H with A and B
// given as padded ELL rows (row-major over slots: slot k of row a at
// base + k*m_l + a).  A or B may be the identity.
struct StencilTerm {
    int32_t in_sector;
    int32_t flags;      // bit0: A identity, bit1: B identity
    int32_t wa, wb;     // ELL widths
    int64_t a_base, b_base;
    double s_re, s_im;
};

struct Stencil {
    std::vector<int64_t> term_base;  // ns+1
    std::vector<StencilTerm> terms;
    std::vector<int32_t> col;        // ELL column indices
    std::vector<cd> val;             // ELL values
    int64_t entries_per_apply = 0;   // sum over outputs of wa*wb
};

// Fused plan of the adjoint Liouvillian under the PI inner product (geometry
// not copied; enough for build_stencil).
void adjoint_plan(const FusedPlan &plan, FusedPlan &adjoint);

// sector_diagonal_only: keep only terms whose input sector is the output
// sector (the sector-block-diagonal part of L, used by the inner solver).
void build_stencil(const FusedPlan &plan, Stencil &st, bool sector_diagonal_only = false);

// sqrt(RN(f)) for the hook-length dimension of a partition (exact big integer).
double sqrt_symmetric_group_dimension(const Partition &p);

// Eigen-decomposition of a complex Hermitian m x m matrix (column-major) by
// cyclic Jacobi rotations: A = U diag(d) U^H.
void hermitian_eigh(int64_t m, std::vector<cd> A, std::vector<double> &d, std::vector<cd> &U);

// Per-sector factors of the stationary-solve preconditioner: U (eigenvectors
// of the fused Hamiltonian block) and inv_den[a,b] = 1/(-i(d_a-d_b) -
// (q_a+q_b)/2), q = diag(U^H Q U); both in the PI block layout (n entries).
// The preconditioner maps R -> U ((U^H R U) .* inv_den) U^H per sector.
void build_sector_preconditioner(const FusedPlan &plan, std::vector<cd> &U, std::vector<cd> &inv_den);

}  // namespace pimi
