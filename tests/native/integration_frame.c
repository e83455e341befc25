/* The renderer-side call sequence of one frame (the one the integration guide shows), written as a C11 caller of
 * include/bzcull.h and linked against libbzcull.so: scene residency, pyramid creation, a transform update through
 * the staging slots, EARLY (frame cache) -> pyramid build -> LATE (frame cache), and the two draw counts.
 *
 *     integration_frame N_OBJECTS
 *
 * Prints "no device" and exits 0 when no GPU is present; otherwise prints "early <count> late <count>" (the caller
 * compares them with the oracle's). Any failing call prints its status and message and exits 1. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bzcull.h"

#define CHECK(call)                                                                                  \
    do {                                                                                             \
        bz_status s_ = (call);                                                                       \
        if (s_ != BZ_OK) {                                                                           \
            fprintf(stderr, "%s: %s (%s)\n", #call, bz_status_str(s_), ctx ? bz_last_error(ctx) : ""); \
            return 1;                                                                                \
        }                                                                                            \
    } while (0)

int main(int argc, char** argv)
{
    const uint32_t n = argc > 1 ? (uint32_t)strtoul(argv[1], NULL, 10) : 100000u;
    const uint64_t seed = 0xB1172E01ull;
    const uint32_t width = 1920, height = 1080;
    bz_ctx* ctx = NULL;
    bz_config cfg;
    memset(&cfg, 0, sizeof cfg);
    const bz_status st = bz_create(&cfg, &ctx);
    if (st == BZ_E_NODEVICE) {
        printf("no device\n");
        return 0;
    }
    CHECK(st);

    /* the renderer's host arrays (here: the seeded synthetic scene) */
    bz_synth_scene_cfg sc;
    memset(&sc, 0, sizeof sc);
    sc.seed = seed;
    sc.n_objects = n;
    sc.n_surfaces = 256;
    sc.lods_per_surface = 8;
    sc.pos_min = -500.0f;
    sc.pos_max = 500.0f;
    sc.scale_min = 0.5f;
    sc.scale_max = 2.0f;
    bz_RenderObject* objs = malloc(sizeof *objs * n);
    bz_MeshTransform* xfs = malloc(sizeof *xfs * n);
    bz_PrimitiveSurface* surfs = malloc(sizeof *surfs * sc.n_surfaces);
    bz_LodData* lods = malloc(sizeof *lods * sc.n_surfaces * sc.lods_per_surface);
    uint32_t* vis = malloc(sizeof *vis * n);
    if (!objs || !xfs || !surfs || !lods || !vis) return 1;
    CHECK(bz_synth_scene(&sc, objs, xfs, surfs, lods));
    CHECK(bz_synth_visibility(seed, n, 0.5f, vis));

    /* once: residency, pyramid, output buffers */
    bz_scene scene = 0;
    CHECK(bz_scene_create(ctx, objs, n, xfs, n, surfs, sc.n_surfaces, lods, sc.n_surfaces * sc.lods_per_surface,
                          &scene));
    CHECK(bz_scene_set_visibility(ctx, scene, vis, n, NULL));
    const uint32_t bw = 1024, bh = 512, mips = 11;
    bz_pyramid pyr = 0;
    CHECK(bz_pyramid_create(ctx, bw, bh, mips, &pyr));
    void *cmds = NULL, *cmds_late = NULL, *counts = NULL, *depth = NULL;
    CHECK(bz_dev_alloc(ctx, (uint64_t)n * sizeof(bz_IndirectDrawData), &cmds));
    CHECK(bz_dev_alloc(ctx, (uint64_t)n * sizeof(bz_IndirectDrawData), &cmds_late));
    CHECK(bz_dev_alloc(ctx, 8, &counts));
    CHECK(bz_dev_alloc(ctx, (uint64_t)width * height * 4, &depth));
    void* stream = NULL;
    CHECK(bz_stream_create(ctx, &stream));

    /* per frame: camera constants, a transform update, the two passes around the pyramid build */
    bz_CameraViewData view;
    const float pos[3] = { 0.0f, 0.0f, 0.0f };
    CHECK(bz_camera_setup(70.0f * (3.14159265358979f / 180.0f), (float)width, (float)height, 0.1f, 650.0f, pos, 0.0f,
                          0.0f, &view));
    view.pyramidWidth = (float)bw;
    view.pyramidHeight = (float)bh;
    uint32_t* up_ids = NULL;
    bz_MeshTransform* up_xf = NULL;
    CHECK(bz_scene_transform_staging(ctx, scene, 4, &up_ids, &up_xf));
    up_ids[0] = 0;
    up_xf[0] = xfs[0]; /* unchanged values: the results stay those of the synthetic scene */
    CHECK(bz_scene_commit_transforms(ctx, scene, 1, stream));

    bz_cull_args a;
    memset(&a, 0, sizeof a);
    a.scene = scene;
    a.view = &view;
    a.cmds_dev = (bz_IndirectDrawData*)cmds;
    a.count_dev = (uint32_t*)counts;
    a.capacity = n;
    a.pass = BZ_PASS_EARLY;
    a.flags = BZ_CULL_FRAME_CACHE;
    CHECK(bz_cull(ctx, &a, stream)); /* early draw list + count */
    /* the renderer draws the early list here and renders depth; the synthetic depth stands in for it */
    CHECK(bz_synth_depth_device(ctx, seed, width, height, 512, view.proj5, view.zNear, (float*)depth, stream));
    CHECK(bz_pyramid_build(ctx, pyr, (const float*)depth, width, height, stream));
    a.pass = BZ_PASS_LATE;
    a.pyramid = pyr;
    a.cmds_dev = (bz_IndirectDrawData*)cmds_late;
    a.count_dev = (uint32_t*)counts + 1;
    CHECK(bz_cull(ctx, &a, stream)); /* newly visible list + count, visibility rewritten in place */
    CHECK(bz_check(ctx, stream));
    uint32_t c[2] = { 0, 0 };
    CHECK(bz_memcpy_d2h(ctx, c, counts, sizeof c, stream));
    printf("early %u late %u\n", c[0], c[1]);

    CHECK(bz_stream_destroy(ctx, stream));
    CHECK(bz_dev_free(ctx, cmds));
    CHECK(bz_dev_free(ctx, cmds_late));
    CHECK(bz_dev_free(ctx, counts));
    CHECK(bz_dev_free(ctx, depth));
    CHECK(bz_pyramid_destroy(ctx, pyr));
    CHECK(bz_scene_destroy(ctx, scene));
    bz_destroy(ctx);
    free(objs);
    free(xfs);
    free(surfs);
    free(lods);
    free(vis);
    return 0;
}
