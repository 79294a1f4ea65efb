// pipeline.h - the phases of one clim_raw_to_rgb frame (internal), shared by
// the single-frame entry points and the burst.
#pragma once

#include "ctx.h"

// A validated clim_raw_to_rgb call and what its phases share.
struct FrameCall {
    const clim_demosaic_params* p;
    int w, h;
    clim::Img out8, outf;           // outputs (outf: null image when not asked for)
    float cam[9], c2y[9], y2c[9];   // rgb_cam, camera -> YCbCr, YCbCr -> camera x exposure
    bool high_noise;                // noiseLevel > 0.6: the raw-domain pre-denoise
    clim::MosaicArg raw;            // the caller's uint16 mosaic
    clim::MosaicArg mosaic;         // the mosaic the demosaic reads (after the pre-denoise: the FP32 plane)
    const float* rawfit;            // device: the raw fit's A[4], B[4] (set by frame_raw_fit)
};

// checks, colour matrices and the context's work buffers for the frame
clim_status frame_prepare(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8, clim_image* out_f32,
                          const clim_demosaic_params* p, FrameCall* fc);
// the context's work buffers as views of the frame's size (no allocation
// once a frame at least as large has been prepared on the context)
clim_status frame_workspace(clim_ctx* ctx, const FrameCall& f);
// raw noise statistics and fit (into st_out, or the context's fit slot 0
// when null), then the high-noise pre-denoise; sets f.mosaic and f.rawfit
// (from f.raw: the phase may run again on the same call)
clim_status frame_raw_fit(clim_ctx* ctx, FrameCall& f, clim::FitState* st_out);
// interpolateGreen into green, interpolateRedBlue + highlights + YCbCr +
// despeckle into ycc
clim_status frame_demosaic(clim_ctx* ctx, const FrameCall& f, const clim::Img& green, const clim::Img& ycc,
                           int rotate_180);
// the pyramid of ycc and the tone map into the frame's outputs
clim_status frame_pyramid(clim_ctx* ctx, const FrameCall& f, const clim::Img& ycc);
