/* fs2d_render.h -- visual output of an fs2d_handle's resident state (C ABI of
 * libfs2d_hip.so).
 *
 * Replaces the kernels of the reference's two renderers:
 *   density_grid_renderer::draw   src/density_grid_renderer.cu:40-56, kernel :10-29
 *   velocity_grid_renderer::draw  src/velocity_grid_renderer.cu:55-72, kernel :8-44
 * The reference copies a host grid to the GPU, converts it there and copies
 * the result back for SFML.  Here the conversion reads the handle's fields in
 * HBM directly (after the steps enqueued before it) and writes either host
 * memory or device memory; drawing the result (an SFML texture or vertex
 * array) stays with the caller.  Results equal the reference kernels' bit for
 * bit, evaluated without FMA contraction.
 *
 * Both cover the whole rows x cols grid, boundary ring included, row-major. */
#ifndef FS2D_RENDER_H
#define FS2D_RENDER_H

#ifdef __cplusplus
extern "C" {
#endif

/* Multipliers of the three color channels (src/density_grid_renderer.cuh:27-31). */
typedef struct fs2d_color_multipliers {
    float r, g, b;
} fs2d_color_multipliers;

/* One vertex of an SFML vertex array (sf::Vertex: position, RGBA color,
 * texture coordinates), 20 bytes, and a line of two
 * (src/velocity_grid_renderer.cuh:11-14). */
typedef struct fs2d_vertex {
    float x, y;
    unsigned char r, g, b, a;
    float tex_x, tex_y;
} fs2d_vertex;

typedef struct fs2d_line {
    fs2d_vertex start, end;
} fs2d_line;

struct fs2d_handle;

/* Density to RGBA8: pixel (i, j) = (c(r d), c(g d), c(b d), 255) with
 * c(p) = (uint8) max(0, min(255, p)) (fmaxf/fminf: NaN -> 255), d the handle's
 * density at (i, j).  rgba: rows*cols*4 bytes of host memory (waits for the
 * result) or, for the _device variant, device memory (enqueued on the
 * handle's stream, returns without waiting). */
int fs2d_hip_render_density(struct fs2d_handle* h, const fs2d_color_multipliers* m, unsigned char* rgba);
int fs2d_hip_render_density_device(struct fs2d_handle* h, const fs2d_color_multipliers* m, unsigned char* rgba);

/* Velocity to lines: line (i, j) starts and ends at (j hs, i vs); where
 * i % 8 == 0 and j % 8 == 0 its end moves by 250000 (u, v) / sqrtf(rows cols).
 * Colors are white, texture coordinates 0.  hs, vs: the reference's
 * target width / cols and target height / rows.  lines: rows*cols lines,
 * host (waits) or device memory (_device, enqueued on the handle's stream). */
int fs2d_hip_render_velocity(struct fs2d_handle* h, float horizontal_scale, float vertical_scale, fs2d_line* lines);
int fs2d_hip_render_velocity_device(struct fs2d_handle* h, float horizontal_scale, float vertical_scale,
                                    fs2d_line* lines);

/* Mean time in milliseconds of each conversion into device memory the handle
 * owns, measured with HIP events on the handle's stream over `reps` runs. */
int fs2d_hip_time_render(struct fs2d_handle* h, int reps, float* ms_density, float* ms_velocity);

#ifdef __cplusplus
}
#endif
#endif
