/*
 * image.h -- BMP texture files and frame dumps (QuadRay core/engine/rtimag.cpp).
 */
#ifndef QR_IMAGE_H
#define QR_IMAGE_H

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace qr {

/* root of the data folders: textures under root + "data/textures/", frame
 * dumps under root + "dump/" (QuadRay's RT_PATH, "../" by default) */
void        set_data_root(const std::string &root);
std::string data_root();

/* a 24-bit uncompressed BMP as texels 0x00RRGGBB, rows as stored */
bool bmp_load(const std::string &path, std::vector<uint32_t> &texels, int &x_dim, int &y_dim);
/* a 24-bit BMP of x_dim x |y_dim| pixels (y_dim < 0: top-down rows), source
 * row y at px + y * stride */
bool bmp_save(const std::string &path, const uint32_t *px, int x_dim, int y_dim,
              ptrdiff_t stride);
/* file name rt_Scene::save_frame gives frame "index" */
std::string frame_name(int index);

}  /* namespace qr */

#endif /* QR_IMAGE_H */
