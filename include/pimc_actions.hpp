// pimc_actions.hpp -- C++ host interface over the pimc_* C ABI, shaped like the
// reference's action interface so a PIMC++ stage can hold these objects in its
// action list.
//
// Reference interfaces mirrored (paths relative to the PIMC++ source tree):
//   ActionBaseClass            PIMC++/src/Actions/ActionBase.h:31-75
//     SingleAction(slice1, slice2, activeParticles, level)   :43
//     Action(...)                                             ActionBase.cc:57-76
//     AcceptCopy / RejectCopy(slice1, slice2)                 :72-73
//     GetName                                                 :71
//     TimeSpent                                               :37
//   MirroredClass::SetMode / ActiveCopy                       PIMC++/src/MirroredClass.h:23-39
//   ShortRangeClass, KineticClass, LongRangeClass             PIMC++/src/Actions/*.cc
//
// Differences from the reference, by design:
//   - the active copy is a member of Path (Path::SetMode), not a process-wide
//     global, so two paths can be evaluated side by side;
//   - errors throw pimc::Error with the library's message where the reference
//     asserts or aborts;
//   - d_dBeta(slice1, slice2, level) evaluates on the path's active copy, as the
//     reference's reads do.
#pragma once

#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

#include "pimc_gpu.h"

namespace pimc {

struct Error : std::runtime_error {
  int code;
  Error(const std::string& what, int c) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char* call) {
  if (rc != PIMC_OK) throw Error(std::string(call) + ": " + pimc_last_error(), rc);
}

enum ModeType { OLDMODE = PIMC_OLDMODE, NEWMODE = PIMC_NEWMODE };

// A context (one device, one stream).  Owns the library objects created through it.
class Context {
 public:
  explicit Context(int device = 0) { check(pimc_ctx_create(device, &ctx_), "pimc_ctx_create"); }
  ~Context() { pimc_ctx_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  pimc_ctx* get() const { return ctx_; }
  void Synchronize() { check(pimc_ctx_synchronize(ctx_), "pimc_ctx_synchronize"); }

 private:
  pimc_ctx* ctx_ = nullptr;
};

// The device-resident path of one replica (the parts of PathClass the actions read),
// with the mirrored-copy selector of MirroredClass.
class Path {
 public:
  Path(Context& ctx, const pimc_replica_desc& desc) : ctx_(ctx), N_(desc.num_particles), M_(desc.num_time_slices) {
    check(pimc_replica_create(ctx.get(), &desc, &id_), "pimc_replica_create");
  }
  int Replica() const { return id_; }
  int NumParticles() const { return N_; }
  int NumTimeSlices() const { return M_ + 1; }   // the reference's NumTimeSlices() counts M+1
  Context& Ctx() const { return ctx_; }

  // MirroredClass::SetMode / the global ActiveCopy.
  void SetMode(ModeType m) { mode_ = m; }
  ModeType GetMode() const { return mode_; }

  // beads [(M+1)][N][3]
  void Set(const std::vector<double>& beads, int mode = PIMC_BOTHMODES) {
    check(pimc_path_set(ctx_.get(), id_, mode, beads.data()), "pimc_path_set");
  }
  std::vector<double> Get(ModeType m) const {
    std::vector<double> b((size_t)(M_ + 1) * N_ * 3);
    check(pimc_path_get(ctx_.get(), id_, m, b.data()), "pimc_path_get");
    return b;
  }
  // Path.SetPos for slices [slice1, slice2] of one particle in the active copy.
  void SetPos(int ptcl, int slice1, int slice2, const std::vector<double>& xyz) {
    check(pimc_path_set_beads(ctx_.get(), id_, mode_, ptcl, slice1, slice2, xyz.data()), "pimc_path_set_beads");
  }
  // PathClass::AcceptCopy / RejectCopy over the moved particles.
  void AcceptCopy(int slice1, int slice2, const std::vector<int>& ptcls) {
    check(pimc_accept_copy(ctx_.get(), id_, slice1, slice2, ptcls.data(), (int)ptcls.size()), "pimc_accept_copy");
  }
  void RejectCopy(int slice1, int slice2, const std::vector<int>& ptcls) {
    check(pimc_reject_copy(ctx_.get(), id_, slice1, slice2, ptcls.data(), (int)ptcls.size()), "pimc_reject_copy");
  }
  void MoveJoin(int old_join, int new_join) {
    check(pimc_move_join(ctx_.get(), id_, old_join, new_join), "pimc_move_join");
  }
  // PathClass::UpdateRho_ks()
  void UpdateRho_ks() { check(pimc_rho_k_update_all(ctx_.get(), id_), "pimc_rho_k_update_all"); }

 private:
  Context& ctx_;
  int id_ = -1, N_ = 0, M_ = 0;
  ModeType mode_ = OLDMODE;
};

// ActionBaseClass: SingleAction on the path's active copy.
class ActionBase {
 public:
  explicit ActionBase(Path& path) : Path_(path) {}
  virtual ~ActionBase() = default;
  virtual double SingleAction(int slice1, int slice2, const std::vector<int>& activeParticles, int level) = 0;
  // ActionBase.cc:57-76: Action forwards to SingleAction.
  virtual double Action(int slice1, int slice2, const std::vector<int>& activeParticles, int level) {
    return SingleAction(slice1, slice2, activeParticles, level);
  }
  // ActionBaseClass::d_dBeta (ActionBase.h:57), the action's beta derivative over links
  // [slice1, slice2) at `level`.
  virtual double d_dBeta(int slice1, int slice2, int level) = 0;
  virtual std::string GetName() = 0;
  // The actions keep no per-move state outside the path (rho_k lives with the path).
  virtual void AcceptCopy(int, int) {}
  virtual void RejectCopy(int, int) {}
  double TimeSpent = 0.0;   // seconds, as ActionBaseClass::TimeSpent

 protected:
  Path& Path_;
  pimc_action_req Request(int slice1, int slice2, const std::vector<int>& ptcls, int level) const {
    if (ptcls.empty() || ptcls.size() > PIMC_MAX_ACTIVE)
      throw Error("SingleAction: 1.." + std::to_string(PIMC_MAX_ACTIVE) + " active particles", PIMC_ERR_ARG);
    pimc_action_req r{};
    r.replica = Path_.Replica();
    r.mode = Path_.GetMode();
    r.slice1 = slice1;
    r.slice2 = slice2;
    r.level = level;
    r.nptcls = (int)ptcls.size();
    for (size_t i = 0; i < ptcls.size(); i++) r.ptcls[i] = ptcls[i];
    return r;
  }
  double DBeta(int32_t action, int slice1, int slice2, int level) const {
    const pimc_range_req r{Path_.Replica(), Path_.GetMode(), slice1, slice2, level, 0};
    double out = 0.0;
    check(pimc_action_d_dbeta(Path_.Ctx().get(), action, &r, 1, &out), "pimc_action_d_dbeta");
    return out;
  }
  template <typename F>
  double Timed(F&& f) {
    const auto t0 = std::chrono::steady_clock::now();
    double v = f();
    TimeSpent += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return v;
  }
};

// ShortRangeClass (PIMC++/src/Actions/ShortRangeClass.cc:131-176).
class ShortRange : public ActionBase {
 public:
  ShortRange(Path& path, int precision = PIMC_FP64) : ActionBase(path), precision_(precision) {}
  double SingleAction(int slice1, int slice2, const std::vector<int>& active, int level) override {
    const pimc_action_req r = Request(slice1, slice2, active, level);
    return Timed([&] {
      double out = 0.0;
      check(pimc_action_short_range(Path_.Ctx().get(), &r, 1, precision_, &out), "pimc_action_short_range");
      return out;
    });
  }
  // ShortRangeClass::d_dBeta (ShortRangeClass.cc:256-295)
  double d_dBeta(int slice1, int slice2, int level) override {
    return DBeta(PIMC_ACTION_SHORT_RANGE, slice1, slice2, level);
  }
  std::string GetName() override { return "ShortRange"; }

 private:
  int precision_;
};

// KineticClass (PIMC++/src/Actions/KineticClass.cc:33-90).
class Kinetic : public ActionBase {
 public:
  using ActionBase::ActionBase;
  double SingleAction(int slice1, int slice2, const std::vector<int>& active, int level) override {
    const pimc_action_req r = Request(slice1, slice2, active, level);
    return Timed([&] {
      double out = 0.0;
      check(pimc_action_kinetic(Path_.Ctx().get(), &r, 1, &out), "pimc_action_kinetic");
      return out;
    });
  }
  // KineticClass::d_dBeta (KineticClass.cc:187-256)
  double d_dBeta(int slice1, int slice2, int level) override { return DBeta(PIMC_ACTION_KINETIC, slice1, slice2, level); }
  std::string GetName() override { return "Kinetic"; }
};

// LongRangeClass (PIMC++/src/Actions/LongRangeClass.cc:273-371): in NEWMODE the
// call first brings rho_k of the NEW copy up to date (PathClass::UpdateRho_ks).
class LongRange : public ActionBase {
 public:
  LongRange(Path& path, int precision = PIMC_FP64) : ActionBase(path), precision_(precision) {}
  double SingleAction(int slice1, int slice2, const std::vector<int>& active, int level) override {
    const pimc_action_req r = Request(slice1, slice2, active, level);
    return Timed([&] {
      double out = 0.0;
      check(pimc_action_long_range(Path_.Ctx().get(), &r, 1, precision_, &out), "pimc_action_long_range");
      return out;
    });
  }
  // LongRangeClass::d_dBeta (LongRangeClass.cc:374-442)
  double d_dBeta(int slice1, int slice2, int level) override {
    return DBeta(PIMC_ACTION_LONG_RANGE, slice1, slice2, level);
  }
  std::string GetName() override { return "LongRange"; }

 private:
  int precision_;
};

// StageClass::StageAction (PIMC++/src/Moves/StageClass.cc:124-136): the sum of a
// stage's actions on the active copy.
inline double StageAction(std::vector<ActionBase*>& actions, int slice1, int slice2,
                          const std::vector<int>& activeParticles, int level) {
  double total = 0.0;
  for (ActionBase* a : actions) total += a->SingleAction(slice1, slice2, activeParticles, level);
  return total;
}

}  // namespace pimc
