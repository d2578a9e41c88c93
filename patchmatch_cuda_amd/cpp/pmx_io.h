/*
 * pmx_io.h -- the reference program's file formats, without OpenCV or boost:
 *   - the INI parameter file of src/main.cpp:42-91 ([params] section);
 *   - Middlebury camera lists (src/image.cpp:64-91) and NVM_V3 models with
 *     per-camera depth ranges (src/image.cpp:93-244);
 *   - 8/16-bit PNG (zlib), binary/ASCII PGM and PPM images, converted to gray
 *     with OpenCV's BGR2GRAY integer weights, as cv::imread + cvtColor do;
 *   - the depth-map file of Array2D_wrapper::saveToFile
 *     (src/array2d_wrapper.h:51-70): "W&H&D&" followed by raw elements.
 */
#ifndef PMX_IO_H
#define PMX_IO_H

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "pmx_dropin.h"

/* key -> value of the [params] section (keys without the section prefix) */
bool pmx_read_ini(const std::string& fileName, std::map<std::string, std::string>& out,
                  std::string* error = nullptr);

/* 8-bit gray image (PNG, PGM/PPM P2/P3/P5/P6); false with a message on error. */
bool pmx_read_gray(const std::string& fileName, GrayImage& out, std::string* error = nullptr);

/* src/image.cpp:64-91: "N" then N lines "name K(9) R(9) T(3)". */
bool readMiddleBurry(std::string fileName, std::vector<Image>& allImages);

/* src/image.cpp:93-244.  The depth range of each camera is the range of
 * the z of the model points it observes, widened by 25% on both sides
 * (the near end shrinking its margin until it stays positive).  The
 * reference takes depthRange by value, so its caller never sees the
 * ranges; here it is filled for the caller. */
bool readNVM(std::string fileName, std::vector<Image>& allImages,
             std::vector<std::pair<float, float>>& depthRange);

/* Reads a file written by saveToFile(name). */
bool pmx_read_depth_file(const std::string& fileName, int& width, int& height, int& depth,
                         std::vector<float>& data);

#endif
