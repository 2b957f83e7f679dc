// fed_hbm.hip — the liveness kernels' device-memory tier.
//
// k_stop and k_gfp keep a node's zone, its scratch zones and its federation lists (stop-set
// escapes and pieces, fixpoint candidates, products and pieces) in LDS, in three launches of
// growing lists. A node whose federations outgrow the largest LDS lists reruns here: the same
// kernels, compiled from the same source with HORA_FED_HBM, which puts the lists in a
// per-workgroup slice of a device buffer (up to 4096 zones per list, the reference's federation
// cap) and makes the wavefront's ordering point a workgroup fence over global memory too.
#define HORA_FED_HBM 1
#include "hora_hip.hip"
