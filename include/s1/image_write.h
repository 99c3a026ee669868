/*
 * image_write.h -- drop-in image writers (image_write.h / image_write.cpp
 * of the reference, and the cf32 dumps of image_write_main.cpp:186-324):
 * the scaled image as a float BigTIFF, last row first, and complex data as
 * raw complex64.
 */
#pragma once

#include "misc_types.h"

/* rows x cols floats as an uncompressed BigTIFF, last data row written
 * first (image_write.cpp:3-25). */
void _write_tif(std::vector<float>& img_data, const int& rows, const int& cols, const std::string out_filename);

/* scale(img_data, scaling_mode), then _write_tif (image_write.cpp:28-40). */
void write_tif(CF_VEC_2D& img_data, const std::string& out_filename, const std::string& scaling_mode);

/* Row-major complex64 samples, each part rounded from double, no header
 * (save_burst_as_cf32 / save_swath_as_cf32, image_write_main.cpp:186-324). */
void write_cf32(const CF_VEC_2D& data, const std::string& out_filename);
