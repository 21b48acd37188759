/*
 * rtm_zfp.h -- this library's codec for ZFP 0.5.5 streams of 2-D float
 * fields (host C).  Internal to librtm_hip.so; the public entry points are
 * declared in include/rtm_hip.h.
 *
 * rtmUq2D stores its seismograms as headerless ZFP 0.5.5 streams
 * (SRC/compression_tools.c:33-70 encode, :94-132 decode): zfp_field_2d over
 * the flat array with nx = dimension1 varying fastest, one of ZFP's lossy
 * modes (the reference uses fixed accuracy), 64-bit stream words.  A stream
 * carries no mode, so encoder and decoder must be given the same one.
 */
#ifndef RTM_ZFP_H
#define RTM_ZFP_H

#include <stddef.h>
#include <stdint.h>

#include "rtm_hip.h"

/* The mode struct and the entry points (rtm_zfp_mode_*, rtm_zfp2f_*) are
 * declared in the public header. */
#endif /* RTM_ZFP_H */
