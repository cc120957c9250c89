// Minimal stand-in for the reference's MGSolT and the members the integration guide's binding touches, with the
// reference's names and types (include/MGSolverBase.h:37-98, include/MGMesh.h:65-80,
// template_appl/nse/SRC/MGSolverT.h:23-110, contrib/matrix/numeric_vectorM.h:160,239).  Test infrastructure: it
// only lets tests/test_integration_binding.py compile and link the binding; it holds no FEMuS code.
#pragma once
#include <vector>

#define DIMENSION 3
#define UP_WIND_T 1.

struct NumericVectorM {
  virtual ~NumericVectorM() {}
  virtual void set(const int i, const double value) = 0;
  virtual double operator()(const int i) const = 0;
};

struct MGMesh {
  int _NoFamFEM = 2;
  int *_NoNodes = nullptr;
  int **_off_el = nullptr;
  int **_el_map = nullptr;
  int **_el_neighbor = nullptr;
  double *_xyz = nullptr;
};

class MGSolT {
 public:
  MGSolT(MGMesh &mesh, int nlev) : _mgmesh(mesh), _NoLevels(nlev) {}
  virtual ~MGSolT() {}
  virtual void bc_read(int, int, double[], int[], int[]) {}
  virtual void GenMatRhs(const double, const int, const int) {}
  virtual void MGSolve(double, int, const int = 1, const int = 8, const int = 40, const int = 8) {}

 protected:
  MGMesh &_mgmesh;
  const int _NoLevels;
  int *_Dim = nullptr;
  std::vector<NumericVectorM *> x, x_old;
  double _dt = 0.1, _IPrdl = 1., _IRe = 1.;
};
