// RCBForceTree.h -- drop-in for the reference's short-range force tree
// (RCBForceTree.h:92-203): the same class template, constructor signature
// and typedefs, implemented over the C ABI of libhaccsr.so (hacc_sr.h) so
// the tree, the interaction lists and the kick run on an MI355X.
//
// As in the reference, the constructor does all the work: it permutes the
// ten caller-owned particle arrays into tree order, builds the lists of the
// leaves inside the force box and adds fcoeff * F to the velocities
// (RCBForceTree.cxx:336-450).  Failures throw std::runtime_error carrying
// hsr_last_error() (the reference asserts or exits instead).
#ifndef RCBForceTree_h
#define RCBForceTree_h

#include <stdint.h>

#include <string>
#include <vector>

#include "ForceLaw.h"

#ifndef BasicDefinition_h
// the reference's build types (BasicDefinition.h:64-101 with -DID_64
// -DPOSVEL_32 -DGRID_32, halo_finder/include.mk:4)
typedef int64_t ID_T;
typedef float POSVEL_T;
typedef uint16_t MASK_T;
#ifndef HACC_SR_DIMENSION_DEFINED
#define HACC_SR_DIMENSION_DEFINED
const int DIMENSION = 3;
#endif
#endif

#define QUADRUPOLE_TDPTS 12
#define MONOPOLE_TDPTS 1

struct hsr_ctx;

// Selects the leaf kernel when a ForceLaw is passed.  The reference's BG/Q
// production build always runs Step16_int (the 5th-order polynomial law,
// BGQStep16.c) whatever ForceLaw is given; its generic build calls
// fl->f_over_r.  Default: the BG/Q (Step16) behaviour.
enum class RCBKernel { Step16, ForceLaw };

template <int TDPTS>
class RCBForceTree {
 public:
  RCBForceTree(POSVEL_T *minLoc, POSVEL_T *maxLoc, POSVEL_T *minForceLoc,
               POSVEL_T *maxForceLoc, ID_T count, POSVEL_T *xLoc,
               POSVEL_T *yLoc, POSVEL_T *zLoc, POSVEL_T *xVel,
               POSVEL_T *yVel, POSVEL_T *zVel, POSVEL_T *mass,
               POSVEL_T *phiLoc, ID_T *idLoc, MASK_T *maskLoc,
               POSVEL_T avgMass, POSVEL_T fsm, POSVEL_T r, POSVEL_T oa,
               ID_T nd = 1, ID_T ds = 1, ID_T tmin = 128, ForceLaw *fl = 0,
               float fcoeff = 0.0, POSVEL_T ppc = 0.9);

  ~RCBForceTree();

  void printStats(double buildTime);

  // process-wide kernel selection for later constructions
  static void setKernel(RCBKernel k);
  // HIP device used by later constructions (default 0)
  static void setDevice(int device);

 protected:
  ID_T particleCount;
  hsr_ctx *ctx;
  double buildTime;

  static RCBKernel s_kernel;
  static int s_device;
};

typedef RCBForceTree<QUADRUPOLE_TDPTS> RCBQuadrupoleForceTree;
typedef RCBForceTree<MONOPOLE_TDPTS> RCBMonopoleForceTree;

#endif  // RCBForceTree_h
