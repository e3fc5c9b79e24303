/*
 * The GRRT C ABI, served by libgrrt.so on one MI355X (gfx950) GPU.
 * Every declaration replaces the one at the cited line of the reference's
 * include/grrt/api.h, with the same signature and meaning.
 */
#ifndef GRRT_API_H
#define GRRT_API_H

#include "grrt_export.h"
#include "grrt/types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* api.h:11 — build a context (metric, camera, disk/star tables uploaded to
 * the GPU).  Returns NULL and sets grrt_last_error() if no GPU is usable. */
GRRT_EXPORT GRRTContext* grrt_create(const GRRTParams* params);
/* api.h:12 */
GRRT_EXPORT void grrt_destroy(GRRTContext* ctx);
/* api.h:13 — copies the parameters and rebuilds the scene from them. */
GRRT_EXPORT void grrt_update_params(GRRTContext* ctx, const GRRTParams* params);
/* api.h:14 */
typedef void (*grrt_progress_fn)(float fraction, void* user_data);

/* api.h:16 — RGBA float32 linear HDR, width*height*4, row 0 = image top. */
GRRT_EXPORT int grrt_render(GRRTContext* ctx, float* framebuffer);
/* api.h:17-18 — progress is called on the caller's thread. */
GRRT_EXPORT int grrt_render_cb(GRRTContext* ctx, float* framebuffer,
                               grrt_progress_fn progress, void* user_data);
/* api.h:19-20 — renders the tile [x,x+w)x[y,y+h) of the full frame into
 * buffer (tile_width*tile_height*4 floats); identical to the same pixels of
 * a full-frame render. */
GRRT_EXPORT int grrt_render_tile(GRRTContext* ctx, float* buffer,
                                 int x, int y, int tile_width, int tile_height);
/* api.h:21 — safe from another thread; the running render stops fetching
 * new rays and returns 1. */
GRRT_EXPORT void grrt_cancel(GRRTContext* ctx);
/* api.h:22 */
GRRT_EXPORT float grrt_progress(const GRRTContext* ctx);
/* api.h:23 */
GRRT_EXPORT const char* grrt_error(const GRRTContext* ctx);
/* api.h:24 — reports whether a GPU (HIP device) is usable. */
GRRT_EXPORT int grrt_cuda_available(void);
/* api.h:25 */
GRRT_EXPORT const char* grrt_last_error(void);

/* api.h:28 */
GRRT_EXPORT void grrt_tonemap(float* framebuffer, int width, int height);

/* api.h:32 — traces one pixel centre and prints a per-ray summary. */
GRRT_EXPORT void grrt_debug_pixel(GRRTContext* ctx, int px, int py);

/* api.h:35-37 */
GRRT_EXPORT void grrt_set_frequency_bins(GRRTContext* ctx,
                                         const double* frequencies_hz,
                                         int num_bins);
/* api.h:38-39 */
GRRT_EXPORT int grrt_render_spectral(GRRTContext* ctx, double* spectral_buffer,
                                     int width, int height);
/* api.h:40-42 */
GRRT_EXPORT int grrt_render_spectral_cb(GRRTContext* ctx, double* spectral_buffer,
                                        int width, int height,
                                        grrt_progress_fn progress, void* user_data);
/* api.h:47-51 */
GRRT_EXPORT int grrt_render_spectral_to_fits_cb(GRRTContext* ctx,
                                                const char* output_path,
                                                int width, int height,
                                                grrt_progress_fn progress,
                                                void* user_data);

#ifdef __cplusplus
}
#endif

#endif
