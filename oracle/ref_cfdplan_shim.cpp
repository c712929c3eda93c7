// ref_cfdplan_shim.cpp -- exports the reference CfdPlan factors (compiled from
// the reference's own bgprtm/Cfd_plan.cpp, which builds standalone) so tests
// can pin the oracle's restatement of it.  Test infrastructure only.
#include "Cfd_plan.h"
#include <cstring>
extern "C" void ref_cfd_plan(int n, float d, float *bm, float *bl, float *am) {
  CfdPlan p(n, d);
  // bm/bl are stored transposed ([i*5+j], [i*4+j]); am is [j*5+i] with 60 entries
  memcpy(bm, p.bm1, sizeof(float) * 5 * n);
  memcpy(bm + 5 * n, p.bm2, sizeof(float) * 5 * n);
  memcpy(bl, p.bl1, sizeof(float) * 4 * n);
  memcpy(bl + 4 * n, p.bl2, sizeof(float) * 4 * n);
  memcpy(am, p.am1, sizeof(float) * 60);
  memcpy(am + 60, p.am2, sizeof(float) * 60);
}
