// Two-site DMRG driver on device-resident tensors (sources/dmrg.cpp).
#pragma once

#include <functional>
#include <limits>
#include <vector>

#include "btensor.hpp"

namespace qtt {

// dmrg_options (include/dmrg_options.h:15-57) with the reference defaults.
struct DmrgOptions {
  double cutoff = 1e-6;
  double convergence_criterion = 1e-5;
  size_t maximum_bond = std::numeric_limits<size_t>::max();
  size_t minimum_bond = 4;
  size_t maximum_iterations = 1000;
};

// Per-sweep record, the dmrg_logger hooks' content (include/dmrg_logger.h:27-185).
struct DmrgLog {
  std::vector<double> energies;   // energy after each sweep
  std::vector<double> sweep_ms;   // wall time of each sweep
  std::vector<int64_t> max_bond;  // largest bond after each sweep
  std::vector<int64_t> middle_bond;
  double contract_flops = 0, svd_flops = 0;
  int64_t updates = 0;
  int64_t lanczos_vectors = 0;  // Krylov vectors built (k-step Lanczos mode)
};

struct DmrgState {
  std::vector<BTensor> mpo;  // W_i (l, s, r, s')
  std::vector<BTensor> mps;  // A_i (l, s, r), orthogonality centre at oc
  std::vector<BTensor> w12;  // two-site Hamiltonians
  std::vector<BTensor> env;  // env[i + 1] = Env[i] for i = -1..L
  int oc = 0;
  int step = 1;
};

// Initialises environments and two-site Hamiltonians (dmrg.cpp:92-100,
// 370-409, 503-515).
void dmrg_prepare(DmrgState& st);
// One call of dmrg_2sites_update::operator() (dmrg.cpp:163-206); returns the
// rank-0 energy tensor.
BTensor dmrg_update(DmrgState& st, const DmrgOptions& opt, DmrgLog* log);
// One full sweep: 2(L-2) updates (dmrg.cpp:127-142); returns the last energy.
double dmrg_sweep(DmrgState& st, const DmrgOptions& opt, DmrgLog* log);
// The calls dmrg_impl makes on its dmrg_logger (sources/dmrg.cpp:239-268):
// init before the first sweep, it_log_all(iteration, E, state) after every
// sweep, end_log_all(iteration, E, state) once the centre is back where it
// started.  Empty functions are skipped.
struct DmrgHooks {
  std::function<void()> init;
  std::function<void(size_t, double, const std::vector<BTensor>&)> it_log_all;
  std::function<void(size_t, double, const std::vector<BTensor>&)> end_log_all;
};
// details::dmrg_impl (dmrg.cpp:219-273) from the orthogonality centre
// st.oc; the centre is moved back there at the end.
double dmrg_run(DmrgState& st, const DmrgOptions& opt, DmrgLog* log, const DmrgHooks* hooks = nullptr);

// Environment helpers (dmrg.cpp:334-345, 424-493).
Shape edge_shape_prep(const Shape& t, int dim);
BTensor compute_left_env(const BTensor& W, const BTensor& Y, const BTensor& L);
BTensor compute_right_env(const BTensor& W, const BTensor& Y, const BTensor& R);
BTensor hamil2site_times_state(const BTensor& theta, const BTensor& W12, const BTensor& L, const BTensor& R);

// one_step_lanczos (sources/dmrg.cpp:584-608): Hpsi is the normalised
// Lanczos vector psi1 = (H theta - a0 theta) / b (not divided when |b| <
// 1e-15), a0 = <theta|H theta>, b = |H theta - a0 theta|, a1 = <psi1|H psi1>,
// all rank-0 tensors on device.
struct LanczosStep {
  BTensor Hpsi, a0, a1, b;
};
LanczosStep one_step_lanczos(const BTensor& theta, const BTensor& W12, const BTensor& L, const BTensor& R,
                             ContractStats* cs = nullptr);
// eig2x2Mat (sources/dmrg.cpp:543-572) on device: the lowest eigenvalue E of
// [[a0, b], [b, a1]] and its eigenvector (o, n).  flag (one int) becomes 1
// when a NaN was met; the caller reads it (the reference throws
// std::logic_error).
struct Eig2x2 {
  BTensor E, o, n;
  BufPtr flag;
};
Eig2x2 eig2x2_dev(const BTensor& a0, const BTensor& a1, const BTensor& b);

// k-step Lanczos with full reorthogonalisation (FP64, on device) for the
// lowest eigenpair of the two-site Hamiltonian L.W.R: stops when the residual
// |H theta' - E theta'| is at most tol |E|, at an invariant subspace, or after
// max_krylov vectors.  With max_krylov = 2 it is the reference's one-step
// update (the 2x2 problem of one_step_lanczos and eig2x2Mat).
struct LanczosResult {
  BTensor theta;  // normalised Ritz vector
  BTensor E;      // rank-0 Ritz value on device
  double energy = 0, residual = 0;
  int iterations = 0;
  int64_t graph_replays = 0;  // matvecs run as a replayed HIP graph
};
LanczosResult lanczos_ground(const BTensor& theta, const BTensor& W12, const BTensor& L, const BTensor& R,
                             int max_krylov, double tol, ContractStats* cs = nullptr);

// bMPS::move_oc (sources/MPT.cpp:74-111): moves the orthogonality centre from
// oc to target with untruncated block SVDs; oc is updated.
void mps_move_oc(std::vector<BTensor>& mps, int& oc, int target);
// The preparation the reference's DMRG drivers apply to a random bMPS
// (tests/DMRG_heisenberg_test.h:34-37): A_0 /= sqrt(<psi|psi>), then the
// orthogonality centre goes to L-1 and back to 0.
void mps_prepare(std::vector<BTensor>& mps);

// <a|b> and <a|H|b> (sources/MPT.cpp:211-292) for checks.
double contract_norm(const std::vector<BTensor>& a, const std::vector<BTensor>& b);
double contract_energy(const std::vector<BTensor>& a, const std::vector<BTensor>& b, const std::vector<BTensor>& mpo);

}  // namespace qtt
