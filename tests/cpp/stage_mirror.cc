// A bisection-stage round trip written against include/pimc_actions.hpp the way a
// PIMC++ stage drives its actions (StageClass::Attempt: SetMode(OLD), StageAction,
// SetMode(NEW), StageAction, then AcceptCopy or RejectCopy).  Linked against any
// library that exports the pimc_* ABI; prints one JSON line.
#include <cmath>
#include <cstdio>
#include <vector>

#include "pimc_actions.hpp"

int main() {
  using namespace pimc;
  Context ctx(0);
  // a smooth synthetic DavidPA table: 3 tau levels, 40 points on [0.2, 2.5], NumTerms 1 (3 columns + potential)
  const int ntau = 3, np = 40, ncol = 4;
  std::vector<double> y((size_t)ntau * np * ncol);
  for (int l = 0; l < ntau; l++)
    for (int i = 0; i < np; i++) {
      const double r = 0.2 + 2.3 * i / (np - 1.0);
      double* row = &y[((size_t)l * np + i) * ncol];
      row[0] = 1.0 / (r * r);
      row[1] = (l + 1) * 0.05 * std::exp(-r);
      row[2] = (l + 1) * 0.01 * std::exp(-2 * r);
      row[3] = (l + 1) * 0.02 * std::exp(-1.5 * r);
    }
  pimc_david_fit_desc fd{PIMC_GRID_LINEAR, np, 0.2, 2.5, ntau, ncol, 1, 0, 0.05, 0.05, y.data()};
  int fit = -1;
  check(pimc_fit_create_david(ctx.get(), &fd, &fit), "pimc_fit_create_david");
  const int N = 8, M = 16;
  const int32_t count[1] = {N};
  const double lambda[1] = {0.5};
  const int32_t pair_fit[1] = {fit};
  pimc_replica_desc rd{};
  rd.num_particles = N;
  rd.num_time_slices = M;
  rd.num_species = 1;
  rd.box[0] = rd.box[1] = rd.box[2] = 5.0;
  rd.periodic[0] = rd.periodic[1] = rd.periodic[2] = 1;
  rd.tau = 0.05;
  rd.species_count = count;
  rd.lambda = lambda;
  rd.pair_fit = pair_fit;
  Path path(ctx, rd);
  std::vector<double> beads((size_t)(M + 1) * N * 3);
  uint32_t s = 12345u;
  auto uni = [&] { s = s * 1664525u + 1013904223u; return (s >> 8) * (1.0 / 16777216.0); };
  std::vector<double> base((size_t)N * 3);
  for (auto& v : base) v = 5.0 * uni() - 2.5;
  for (int t = 0; t <= M; t++)
    for (int p = 0; p < N; p++)
      for (int d = 0; d < 3; d++) beads[((size_t)t * N + p) * 3 + d] = base[p * 3 + d] + 0.1 * (uni() - 0.5);
  path.Set(beads);

  ShortRange sr(path);
  Kinetic kin(path);
  std::vector<ActionBase*> actions{&kin, &sr};
  const std::vector<int> moved{3};
  const int slice1 = 4, slice2 = 12, level = 1;

  // move particle 3 on slices 5..11 in the NEW copy
  std::vector<double> xyz((size_t)(slice2 - slice1 + 1) * 3);
  for (int t = slice1; t <= slice2; t++)
    for (int d = 0; d < 3; d++) {
      const double shift = (t > slice1 && t < slice2) ? 0.2 * std::sin(0.7 * t + d) : 0.0;
      xyz[(size_t)(t - slice1) * 3 + d] = beads[((size_t)t * N + 3) * 3 + d] + shift;
    }
  path.SetMode(NEWMODE);
  path.SetPos(3, slice1, slice2, xyz);

  path.SetMode(OLDMODE);
  const double s_old = StageAction(actions, slice1, slice2, moved, level);
  path.SetMode(NEWMODE);
  const double s_new = StageAction(actions, slice1, slice2, moved, level);

  // reject: NEW <- OLD, so the NEW action equals the OLD one again
  path.RejectCopy(slice1, slice2, moved);
  const double s_new_after_reject = StageAction(actions, slice1, slice2, moved, level);

  // move again and accept: OLD <- NEW, so the OLD action equals the NEW one
  path.SetPos(3, slice1, slice2, xyz);
  path.AcceptCopy(slice1, slice2, moved);
  path.SetMode(OLDMODE);
  const double s_old_after_accept = StageAction(actions, slice1, slice2, moved, level);

  bool threw = false;
  try {
    sr.SingleAction(slice1, slice2, {}, level);
  } catch (const Error& e) {
    threw = e.code == PIMC_ERR_ARG;
  }
  // d_dBeta through the ActionBase interface: the kinetic one needs no tables; the
  // short-range one needs the fit's beta-derivative table, which this fit lacks
  ActionBase* as_base = &kin;
  const double kin_dbeta = as_base->d_dBeta(0, M, 0);
  bool dbeta_needs_tables = false;
  try {
    sr.d_dBeta(0, M, 0);
  } catch (const Error& e) {
    dbeta_needs_tables = e.code == PIMC_ERR_STATE;
  }
  std::printf("{\"s_old\": %.17g, \"s_new\": %.17g, \"s_new_after_reject\": %.17g, "
              "\"s_old_after_accept\": %.17g, \"empty_request_throws\": %s, \"time_spent\": %.6g, "
              "\"kin_dbeta\": %.17g, \"dbeta_needs_tables\": %s}\n",
              s_old, s_new, s_new_after_reject, s_old_after_accept, threw ? "true" : "false", sr.TimeSpent,
              kin_dbeta, dbeta_needs_tables ? "true" : "false");
  return 0;
}
