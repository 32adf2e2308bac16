// raff_internal.h -- host object behind the smr_raff handle of
// include/surfmol_raff.h, shared by raff.cpp and raff_server.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/surfmol_raff.h"
#include "smr_internal.h"

struct smr_raff {
    smr_ctx *ctx = nullptr;
    int n = 0, nsys = 0, nnode = -1, seg = 0;
    bool collide = false;
    smr::DevBuf radius, neighs, excl1, excl2, bk_slots, fixmask, port_local, kflat;
    smr::DevBuf pos, quat, dpos_mom, dquat_mom, dpos_coll;
    smr::DevBuf dpos_node, drot_node, dpos_neigh;
    smr::DevBuf metric, steps, status;
    std::vector<int> bk_host, neigh_host;
    bool have_state = false, have_neighs = false, have_bk = false;
};


namespace smr {
// segment (lanes per replica) of the persistent kernel for n atoms
int raff_seg_for(int n);
int raff_check_cfg(const smr_raff_step_config *c, const char *who);
// the uploads a step needs are present and consistent
int raff_ready(const smr_raff *r, const char *who);
// kernel arguments of `nsteps` steps over every replica of r
RaffArgs raff_args(smr_raff *r, const smr_raff_step_config *c, int nsteps);
}  // namespace smr
