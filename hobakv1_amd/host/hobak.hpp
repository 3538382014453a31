// Reference-shaped C++ host façade over the C ABI of libhobak_hip.so (include/hobak_c.h).
//
// A caller of the reference's backward-Euler path keeps its shape (paths relative to the
// reference root):
//
//   TET_MESH(restVertices, tets)                  src/Geometry/TET_MESH.h:37-38
//   VOLUME::SNH(mu, lambda), VOLUME::ARAP(mu, lambda)
//                                                 src/Hyperelastic/Volume/SNH.h:28, ARAP.h:28
//   HYPERELASTIC::computeMu / computeLambda       src/Hyperelastic/Volume/HYPERELASTIC.cpp:111-122
//   CUBE(center, scale)                           src/Geometry/CUBE.cpp:27-55
//   BACKWARD_EULER_POSITION(mesh, material)       src/Timestepper/BACKWARD_EULER_POSITION.h:36
//   BACKWARD_EULER_VELOCITY(mesh, material)       src/Timestepper/BACKWARD_EULER_VELOCITY.h:35
//     solve(verbose)                              src/Timestepper/TIMESTEPPER.h:77
//     position(), velocity()                      src/Timestepper/TIMESTEPPER.h:44-58
//     addGravity, attachKinematicConstraints, addKinematicCollisionObject, setDt, setRayeligh,
//     scaleMass; TET_MESH::readTobjFile / writeTobjFile
//                                                 src/Timestepper/TIMESTEPPER.h:70-100
//
// position() is the displacement u = x - X, as in the reference.  The device instance is created
// at the first solve(), after the setters have run (the reference reads its members inside
// solveRayleighDamped).  position() and velocity() return host vectors the caller may edit
// between solves: solve() uploads them, steps on the GPU and downloads the new state.  All
// compute runs in libhobak_hip.so; a missing library or device is an error (hobak::Error).
#pragma once

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hobak_c.h"

namespace hobak {

typedef std::array<double, 3> VECTOR3;
typedef std::array<int, 4> VECTOR4I;
typedef std::vector<double> VECTOR;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

// One device, one stream, an HBM budget (0 = none).
class Context {
 public:
  explicit Context(int device = 0, size_t budget_bytes = 0) {
    const int rc = hobak_ctx_create(device, budget_bytes, &ctx_);
    if (rc != HOBAK_OK) throw Error(rc, "hobak_ctx_create failed (no gfx950 device?)");
  }
  ~Context() { hobak_ctx_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  hobak_ctx* get() const { return ctx_; }
  void check(int rc, const char* what) const {
    if (rc != HOBAK_OK) throw Error(rc, std::string(what) + ": " + hobak_last_error(ctx_));
  }
  static Context& global() {
    static Context c;
    return c;
  }

 private:
  hobak_ctx* ctx_ = nullptr;
};

class TET_MESH {
 public:
  TET_MESH(const std::vector<VECTOR3>& restVertices, const std::vector<VECTOR4I>& tets)
      : rest_(restVertices), tets_(tets) {}
  const std::vector<VECTOR3>& restVertices() const { return rest_; }
  const std::vector<VECTOR4I>& tets() const { return tets_; }
  int DOFs() const { return 3 * (int)rest_.size(); }
  // TET_MESH::computeFdots (TET_MESH.cpp:660-684): the velocity gradients an energy-damped step
  // reads.  An integrator on this mesh takes them to the device at its next solve().
  void computeFdots(const VECTOR& velocity) {
    fdotVelocity_ = velocity;
    fdotSerial_++;
  }
  const VECTOR& fdotVelocity() const { return fdotVelocity_; }
  long fdotSerial() const { return fdotSerial_; }

  // TET_MESH::readTobjFile (TET_MESH.cpp:1173-1220): a run of "v x y z" lines, then a run of
  // "t a b c d" lines; false if the file cannot be opened or ends right after its vertices.
  static bool readTobjFile(const std::string& filename, std::vector<VECTOR3>& vertices,
                           std::vector<VECTOR4I>& tets) {
    vertices.clear();
    tets.clear();
    FILE* f = std::fopen(filename.c_str(), "r");
    if (!f) return false;
    char line[512];
    bool in_tets = false, any_after = false;
    while (std::fgets(line, sizeof(line), f)) {
      if (!in_tets && line[0] == 'v') {
        VECTOR3 v;
        if (std::sscanf(line, "v %lf %lf %lf", &v[0], &v[1], &v[2]) != 3) break;
        vertices.push_back(v);
        continue;
      }
      any_after = any_after || line[0] != '\n';
      if (line[0] != 't') break;
      in_tets = true;
      VECTOR4I t;
      if (std::sscanf(line, "t %d %d %d %d", &t[0], &t[1], &t[2], &t[3]) != 4) break;
      tets.push_back(t);
    }
    std::fclose(f);
    return any_after;
  }
  // TET_MESH::writeTobjFile (TET_MESH.cpp:1225-1259) of the rest vertices: %.17g, so they read
  // back bit for bit.
  static bool writeTobjFile(const std::string& filename, const TET_MESH& mesh) {
    FILE* f = std::fopen(filename.c_str(), "w");
    if (!f) return false;
    for (const auto& v : mesh.restVertices()) std::fprintf(f, "v %.17g %.17g %.17g\n", v[0], v[1], v[2]);
    for (const auto& t : mesh.tets()) std::fprintf(f, "t %i %i %i %i\n", t[0], t[1], t[2], t[3]);
    std::fclose(f);
    return true;
  }

 private:
  std::vector<VECTOR3> rest_;
  std::vector<VECTOR4I> tets_;
  VECTOR fdotVelocity_;
  long fdotSerial_ = 0;
};

namespace VOLUME {
class HYPERELASTIC {
 public:
  HYPERELASTIC(double mu, double lambda) : mu_(mu), lambda_(lambda) {}
  virtual ~HYPERELASTIC() = default;
  virtual std::string name() const = 0;
  virtual int kind() const = 0;
  double mu() const { return mu_; }
  double lambda() const { return lambda_; }
  static double computeMu(double E, double nu) { return E / (2.0 * (1.0 + nu)); }
  static double computeLambda(double E, double nu) { return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)); }

 private:
  double mu_, lambda_;
};
class SNH : public HYPERELASTIC {
 public:
  using HYPERELASTIC::HYPERELASTIC;
  std::string name() const override { return "SNH"; }
  int kind() const override { return HOBAK_MAT_SNH; }
};
class ARAP : public HYPERELASTIC {
 public:
  using HYPERELASTIC::HYPERELASTIC;
  std::string name() const override { return "ARAP"; }
  int kind() const override { return HOBAK_MAT_ARAP; }
};
// VOLUME::GREEN_DAMPING(mu) (src/Damping/Volume/GREEN_DAMPING.cpp:25-28).
class GREEN_DAMPING {
 public:
  explicit GREEN_DAMPING(double mu) : mu_(mu) {}
  std::string name() const { return "GREEN_DAMPING"; }
  double mu() const { return mu_; }

 private:
  double mu_;
};
}  // namespace VOLUME

// Axis-aligned kinematic cube: local = scale^-1 (world - center), world = scale local + center
// (the transforms of src/Geometry/KINEMATIC_SHAPE.h:55-73, rotation = I, multiplying by the
// inverse scale as the reference does, which takes it from Eigen's 3x3 cofactor inverse,
// CUBE.cpp:28); inside when every local coordinate is within [-0.5, 0.5] (CUBE.cpp:44-55).
class CUBE {
 public:
  CUBE(const VECTOR3& center, double scale)
      : c_(center), s_(scale), inv_((scale * scale) * (1.0 / ((scale * scale) * scale))) {}
  const VECTOR3& center() const { return c_; }
  double scale() const { return s_; }
  VECTOR3 local(const VECTOR3& w) const {
    return {inv_ * (w[0] - c_[0]), inv_ * (w[1] - c_[1]), inv_ * (w[2] - c_[2])};
  }
  VECTOR3 world(const VECTOR3& l) const {
    return {s_ * l[0] + c_[0], s_ * l[1] + c_[1], s_ * l[2] + c_[2]};
  }
  bool inside(const VECTOR3& w) const {
    const VECTOR3 l = local(w);
    for (int k = 0; k < 3; k++)
      if (!(l[k] <= 0.5 && l[k] >= -0.5)) return false;
    return true;
  }

 private:
  VECTOR3 c_;
  double s_, inv_;
};

class BACKWARD_EULER_POSITION {
 public:
  BACKWARD_EULER_POSITION(TET_MESH& mesh, VOLUME::HYPERELASTIC& material, Context* ctx = nullptr)
      : BACKWARD_EULER_POSITION(mesh, material, ctx, HOBAK_INTEGRATOR_POSITION) {}
  ~BACKWARD_EULER_POSITION() { hobak_inst_destroy(inst_); }
  BACKWARD_EULER_POSITION(const BACKWARD_EULER_POSITION&) = delete;
  BACKWARD_EULER_POSITION& operator=(const BACKWARD_EULER_POSITION&) = delete;

  VECTOR& position() { return position_; }
  VECTOR& velocity() { return velocity_; }
  void setDt(double dt) { prm_.dt = dt; }
  void setRayeligh(double alpha, double beta) {
    prm_.rayleigh_alpha = alpha;
    prm_.rayleigh_beta = beta;
  }
  void scaleMass(double s) { prm_.mass_scale *= s; }
  void addGravity(const VECTOR3& g) {
    for (int c = 0; c < 3; c++) prm_.gravity[c] = g[c];
  }
  // Not in the reference's public surface (its CG uses Eigen's defaults): the PCG tolerance,
  // the preconditioner and the precision form.
  void setTolerance(double tol) { prm_.pcg_tol = tol; }
  void setPreconditioner(int precond) { prm_.precond = precond; }
  void setPrecision(int precision) { prm_.precision = precision; }
  void setNewtonIterations(int n, double tol = 1e-7) {
    prm_.newton_iters = n;
    prm_.newton_tol = tol;
  }
  void attachKinematicConstraints(const CUBE* shape) {
    const auto& X = mesh_.restVertices();
    for (size_t i = 0; i < X.size(); i++) {
      if (!shape->inside(X[i])) continue;
      bool have = false;
      for (int p : pins_) have = have || p == (int)i;
      if (have) continue;
      pins_.push_back((int)i);
      const VECTOR3 w = shape->world(shape->local(X[i]));
      for (int c = 0; c < 3; c++) targets_.push_back(w[c] - X[i][c]);
    }
  }

  // TIMESTEPPER::addKinematicCollisionObject (TIMESTEPPER.cpp:433-441): surface vertices that
  // enter the cube become plane constraints (FP64 form, one Newton iteration).
  void addKinematicCollisionObject(const CUBE* shape) {
    for (const CUBE* c : collision_)
      if (c == shape) return;
    collision_.push_back(shape);
    if (inst_) add_cube(shape);
  }
  // The integrator's damping object (TIMESTEPPER.h:155; the reference sets it in its constructor,
  // BACKWARD_EULER_POSITION.cpp:71-76): non-null makes solve() take solveEnergyDamped
  // (BACKWARD_EULER_POSITION.cpp:125-131).
  void setDamping(const VOLUME::GREEN_DAMPING* damping) {
    damping_ = damping;
    if (inst_) apply_damping();
  }
  // Vertex ids of the live plane constraints (TIMESTEPPER.h:46), ascending.
  std::vector<int> planeConstraints() const {
    std::vector<int> ids;
    if (!inst_) return ids;
    int32_t n = 0;
    ctx_->check(hobak_inst_plane_constraints(inst_, &n, nullptr, 0), "plane_constraints");
    std::vector<int32_t> buf(n);
    ctx_->check(hobak_inst_plane_constraints(inst_, &n, buf.data(), n), "plane_constraints");
    ids.assign(buf.begin(), buf.end());
    return ids;
  }

  bool solve(bool verbose) {
    ensure();
    ctx_->check(hobak_inst_set_state(inst_, position_.data(), velocity_.data()), "set_state");
    if (damping_ && mesh_.fdotSerial() != fdotApplied_) {
      ctx_->check(hobak_inst_compute_fdots(inst_, mesh_.fdotVelocity().data()), "compute_fdots");
      fdotApplied_ = mesh_.fdotSerial();
    }
    const int rc = hobak_step(inst_, &stats_);
    if (rc != HOBAK_OK && rc != HOBAK_E_NOT_CONVERGED) ctx_->check(rc, "hobak_step");
    ctx_->check(hobak_inst_get_state(inst_, position_.data(), velocity_.data()), "get_state");
    if (verbose) std::printf("  PCG iters: %d err: %.4e\n", stats_.pcg_iters[0], stats_.pcg_error[0]);
    return true;
  }
  const hobak_step_stats& lastStats() const { return stats_; }

 protected:
  BACKWARD_EULER_POSITION(TET_MESH& mesh, VOLUME::HYPERELASTIC& material, Context* ctx, int integrator)
      : mesh_(mesh), material_(material), ctx_(ctx ? ctx : &Context::global()) {
    hobak_default_params(&prm_);
    prm_.precond = HOBAK_PRECOND_BLOCK_JACOBI;
    prm_.integrator = integrator;
    position_.assign(mesh.DOFs(), 0.0);
    velocity_.assign(mesh.DOFs(), 0.0);
  }

 private:
  void ensure() {
    if (inst_) return;
    std::vector<double> rest;
    for (const auto& v : mesh_.restVertices()) rest.insert(rest.end(), v.begin(), v.end());
    std::vector<int32_t> tets;
    for (const auto& t : mesh_.tets()) tets.insert(tets.end(), t.begin(), t.end());
    const hobak_material m{material_.kind(), material_.mu(), material_.lambda()};
    ctx_->check(hobak_inst_create(ctx_->get(), rest.data(), (int32_t)mesh_.restVertices().size(),
                                  tets.data(), (int32_t)mesh_.tets().size(), &m, &prm_, &inst_),
                "hobak_inst_create");
    // pinned vertex ids ascending, targets in the same order
    std::vector<size_t> order(pins_.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = k;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return pins_[a] < pins_[b]; });
    std::vector<int32_t> ids;
    std::vector<double> tg;
    for (size_t k : order) {
      ids.push_back(pins_[k]);
      for (int c = 0; c < 3; c++) tg.push_back(targets_[3 * k + c]);
    }
    if (!ids.empty())
      ctx_->check(hobak_inst_set_pinned(inst_, ids.data(), (int32_t)ids.size(), tg.data()), "set_pinned");
    for (const CUBE* c : collision_) add_cube(c);
    if (damping_) apply_damping();
  }
  void apply_damping() {
    ctx_->check(hobak_inst_set_damping(inst_, damping_ ? HOBAK_DAMPING_GREEN : HOBAK_DAMPING_RAYLEIGH,
                                       damping_ ? damping_->mu() : 0.0),
                "set_damping");
  }
  void add_cube(const CUBE* c) {
    const double center[3] = {c->center()[0], c->center()[1], c->center()[2]};
    ctx_->check(hobak_inst_add_collision_cube(inst_, center, c->scale()), "add_collision_cube");
  }

  TET_MESH& mesh_;
  VOLUME::HYPERELASTIC& material_;
  Context* ctx_;
  hobak_step_params prm_{};
  hobak_inst* inst_ = nullptr;
  hobak_step_stats stats_{};
  VECTOR position_, velocity_;
  std::vector<int> pins_;
  std::vector<double> targets_;
  std::vector<const CUBE*> collision_;
  const VOLUME::GREEN_DAMPING* damping_ = nullptr;
  long fdotApplied_ = 0;
};

// BACKWARD_EULER_VELOCITY(mesh, material) (src/Timestepper/BACKWARD_EULER_VELOCITY.h:35, .cpp:31-46):
// the same surface, solving for the velocity change (BACKWARD_EULER_VELOCITY.cpp:107-235).
class BACKWARD_EULER_VELOCITY : public BACKWARD_EULER_POSITION {
 public:
  BACKWARD_EULER_VELOCITY(TET_MESH& mesh, VOLUME::HYPERELASTIC& material, Context* ctx = nullptr)
      : BACKWARD_EULER_POSITION(mesh, material, ctx, HOBAK_INTEGRATOR_VELOCITY) {}
};

}  // namespace hobak
