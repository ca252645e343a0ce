/*
 * shim_demo.cpp -- drives QuadRay's rt_Scene API as qr_scene.hpp offers it:
 * platform, scene over a caller framebuffer, the locking rule, render and
 * the error path of a platform without a device.
 *
 * usage: shim_demo host
 *        shim_demo gpu <out.bin>    (200x120, 4X, depth 4, plain scene seed 1)
 *        shim_demo camera <out.bin> (200x120, no AA, depth 4, plain scene seed
 *                                    1, camera moved and turned between renders)
 * Exit status 0 on success; the frame is written as raw uint32 rows.
 */
#include <stdio.h>
#include <string.h>

#include <vector>

#include "qr_scene.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const bool camera = strcmp(argv[1], "camera") == 0;
    const bool gpu = camera || strcmp(argv[1], "gpu") == 0;
    int info[5];
    struct rt_SCENE *sc = qrgen_create(1, 1, 0, info);
    if (sc == nullptr) return 2;
    const int x_res = 200, y_res = 120, x_row = 256;
    std::vector<uint32_t> frame((size_t)x_row * y_res, 0u);
    int rc = 0;
    try
    {
        qr::Platform pfm(gpu ? 0 : QRHIP_HOST_ONLY);
        if (pfm.set_fsaa(9) != QRHIP_FSAA_4X) rc = 10;
        if (camera) pfm.set_fsaa(QRHIP_FSAA_NO);
        qr::Scene scn(sc, x_res, y_res, x_row, frame.data(), &pfm);
        if (scn.get_x_row() != x_row) rc = 11;
        if (scn.get_depth() != 10) rc = 12;
        scn.set_depth(4);
        try
        {
            qr::Scene again(sc, x_res, y_res, x_row, nullptr, &pfm);
            rc = 13;
        }
        catch (const qr::Exception &e)
        {
            if (strstr(e.err, "locked") == nullptr) rc = 14;
        }
        if (camera)
        {
            /* RooT's camera keys (root/RooT.h:539-547) as three frames */
            scn.render(0);
            scn.update(100, RT_CAMERA_MOVE_FORWARD);
            scn.update(100, RT_CAMERA_ROTATE_LEFT);
            scn.render(100);
            scn.update(250, RT_CAMERA_MOVE_LEFT);
            scn.update(250, RT_CAMERA_ROTATE_DOWN);
            scn.render(250);
            scn.update(400, RT_CAMERA_MOVE_UP);
            scn.update(400, RT_CAMERA_MOVE_RIGHT);
            scn.update(400, RT_CAMERA_ROTATE_RIGHT);
        }
        if (gpu)
        {
            scn.render(camera ? 400 : 0);
            if (scn.get_frame() != frame.data()) rc = 15;
            FILE *f = argc > 2 ? fopen(argv[2], "wb") : nullptr;
            if (f == nullptr) return 3;
            for (int y = 0; y < y_res; y++)
            {
                fwrite(frame.data() + (size_t)y * x_row, 4, x_res, f);
            }
            fclose(f);
        }
        else
        {
            scn.update(100, RT_CAMERA_MOVE_FORWARD);    /* host side: no device needed */
            try
            {
                scn.render(0);
                rc = 16;
            }
            catch (const qr::Exception &e)
            {
                if (strstr(e.err, "no device") == nullptr) rc = 17;
            }
        }
    }
    catch (const qr::Exception &e)
    {
        fprintf(stderr, "qr::Exception: %s\n", e.err);
        rc = 1;
    }
    qrgen_destroy(sc);
    if (rc != 0) fprintf(stderr, "shim_demo: check %d failed\n", rc);
    return rc;
}
