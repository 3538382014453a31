// Test driver of the C++ façade (hobak.hpp): runs a scene the way SIMULATION_SCENE::stepSimulation
// does (src/Scenes/SIMULATION_SCENE.h:120-126) through TET_MESH / SNH / ARAP / CUBE /
// BACKWARD_EULER_POSITION, and writes the final state and per-step PCG iterations.
//
//   facade_driver <scene.bin> <out.bin>        (HOBAK_FACADE_PRECOND=0: Jacobi PCG)
//
// scene.bin (little-endian): int32 nv, nt, material, steps, flags (1: pin cube, 2: collision
// cube, 4: GREEN_DAMPING, 8: BACKWARD_EULER_VELOCITY in place of BACKWARD_EULER_POSITION); float64 mu, lambda, dt, rayleigh alpha, beta, mass scale, gravity[3],
// pcg tol, pin cube center[3] and scale, collision cube center[3] and scale; then rest (nv*3
// float64), tets (nt*4 int32), u0 and v0 (3*nv float64 each); with flag 4 also the damping mu
// (float64) and the velocity given to TET_MESH::computeFdots before the first step (3*nv float64).
// out.bin: int32 steps, iterations[steps], plane-constraint counts[steps]; float64 u[3*nv],
// v[3*nv].
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "hobak.hpp"

namespace {
template <typename T>
bool rd(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: facade_driver <scene.bin> <out.bin>\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[5];
  double prm[18];
  if (!rd(f, hdr, 5) || !rd(f, prm, 18)) return 2;
  const int nv = hdr[0], nt = hdr[1], mat = hdr[2], steps = hdr[3], flags = hdr[4];
  std::vector<hobak::VECTOR3> rest(nv);
  std::vector<hobak::VECTOR4I> tets(nt);
  for (auto& v : rest)
    if (!rd(f, v.data(), 3)) return 2;
  for (auto& t : tets)
    if (!rd(f, t.data(), 4)) return 2;
  std::vector<double> u0(3 * (size_t)nv), v0(3 * (size_t)nv), vf(3 * (size_t)nv);
  if (!rd(f, u0.data(), u0.size()) || !rd(f, v0.data(), v0.size())) return 2;
  double dmu = 0.0;
  if ((flags & 4) && (!rd(f, &dmu, 1) || !rd(f, vf.data(), vf.size()))) return 2;
  std::fclose(f);

  try {
    hobak::TET_MESH mesh(rest, tets);
    hobak::VOLUME::SNH snh(prm[0], prm[1]);
    hobak::VOLUME::ARAP arap(prm[0], prm[1]);
    hobak::VOLUME::HYPERELASTIC& material = mat == HOBAK_MAT_SNH ? (hobak::VOLUME::HYPERELASTIC&)snh
                                                                  : (hobak::VOLUME::HYPERELASTIC&)arap;
    std::unique_ptr<hobak::BACKWARD_EULER_POSITION> integrator;
    if (flags & 8)
      integrator.reset(new hobak::BACKWARD_EULER_VELOCITY(mesh, material));
    else
      integrator.reset(new hobak::BACKWARD_EULER_POSITION(mesh, material));
    hobak::BACKWARD_EULER_POSITION& solver = *integrator;
    solver.setDt(prm[2]);
    solver.setRayeligh(prm[3], prm[4]);
    solver.scaleMass(prm[5]);
    solver.addGravity({prm[6], prm[7], prm[8]});
    solver.setTolerance(prm[9]);
    // HOBAK_FACADE_PRECOND=0 selects Jacobi, the reference's preconditioner (default: block Jacobi)
    if (const char* pc = std::getenv("HOBAK_FACADE_PRECOND")) solver.setPreconditioner(std::atoi(pc));
    hobak::CUBE cube({prm[10], prm[11], prm[12]}, prm[13]);
    if (flags & 1) solver.attachKinematicConstraints(&cube);
    hobak::CUBE ground({prm[14], prm[15], prm[16]}, prm[17]);
    if (flags & 2) solver.addKinematicCollisionObject(&ground);
    hobak::VOLUME::GREEN_DAMPING damping(dmu);
    if (flags & 4) {
      solver.setDamping(&damping);
      mesh.computeFdots(vf);
    }
    solver.position() = u0;
    solver.velocity() = v0;
    std::vector<int32_t> iters, planes;
    for (int s = 0; s < steps; s++) {
      solver.solve(false);
      iters.push_back(solver.lastStats().pcg_iters[0]);
      planes.push_back((int32_t)solver.planeConstraints().size());
    }
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    const int32_t n = steps;
    std::fwrite(&n, sizeof(n), 1, o);
    std::fwrite(iters.data(), sizeof(int32_t), iters.size(), o);
    std::fwrite(planes.data(), sizeof(int32_t), planes.size(), o);
    std::fwrite(solver.position().data(), sizeof(double), solver.position().size(), o);
    std::fwrite(solver.velocity().data(), sizeof(double), solver.velocity().size(), o);
    std::fclose(o);
  } catch (const hobak::Error& e) {
    std::fprintf(stderr, "facade_driver: %s (code %d)\n", e.what(), e.code);
    return 1;
  }
  return 0;
}
