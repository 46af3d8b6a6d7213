// Internal interface between the network (mcl_net.cpp) and the solver context
// (mcl_api.cpp).  The public entry points are declared in include/mcl.h.
#pragma once
#include <hip/hip_runtime.h>

#include "mcl.h"

struct mcl_net;

// device buffers of the network: input [B][H][W][2], output [B][H][W][3], target
float *mcl_net_in(mcl_net *n);
float *mcl_net_out(mcl_net *n);
float *mcl_net_tgt(mcl_net *n);
hipStream_t mcl_net_stream(mcl_net *n);
int mcl_net_dims(const mcl_net *n, int *H, int *W, int *B);
int mcl_net_forward_device(mcl_net *n);
float mcl_net_out_scale(const mcl_net *n, int channel);
// append one snapshot (device arrays, copied) to the dataset
int mcl_net_add_snapshot(mcl_net *n, const float *x_dev, const float *t_dev);
