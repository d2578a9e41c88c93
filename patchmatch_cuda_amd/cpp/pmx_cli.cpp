/*
 * pmx_cli -- the reference program (src/main.cpp:10-128) on the MI355X
 * engine: reads the INI parameter file, the camera list (NVM_V3 or
 * Middlebury), runs PatchMatch for the reference view and writes the depth
 * map with Array2D_wrapper::saveToFile.
 *
 *   pmx_cli <config.ini>
 *
 * [params] keys (src/main.cpp:47-85): imageInfoFileName, depthRangeNear,
 * depthRangeFar, gpuId, orientationX, orientationZ, refImageId,
 * halfWindowSize, numOfSamples, SPMAlpha, gaussianSigma, numOfIterations,
 * outputFileName.  Engine keys with no reference counterpart: blurMode
 * (none | gauss | box, default gauss), initMode (depth | inverse, default
 * depth), rngSeed (default 7), outputCostFileName, outputIdsFileName.
 */
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "pmx_dropin.h"
#include "pmx_io.h"

namespace {

struct Ini {
    std::map<std::string, std::string> kv;
    bool has(const char* k) const { return kv.count(k) != 0; }
    std::string str(const char* k) const {
        auto it = kv.find(k);
        if (it == kv.end()) throw std::runtime_error(std::string("No such node (params.") + k + ")");
        return it->second;
    }
    float f(const char* k) const { return std::stof(str(k)); }
    int i(const char* k) const { return std::stoi(str(k)); }
};

} // namespace

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cout << "Initialization file is mandatory to run the code" << std::endl;
        return 0;
    }
    Ini ini;
    std::string err;
    std::string imageInfoFileName, outputFileName;
    float nearRange = -1.0f, farRange = -1.0f, orientationX = 0.0f, orientationZ = 1.0f;
    float SPMAlpha = 0.0f, gaussianSigma = 0.0f;
    int refImageId = 0, halfWindowSize = 0, numOfSamples = 0, numOfIterations = 0, gpuId = -1;
    try {
        if (!pmx_read_ini(argv[1], ini.kv, &err)) throw std::runtime_error(err);
        imageInfoFileName = ini.str("imageInfoFileName");
        if (ini.has("depthRangeNear")) nearRange = ini.f("depthRangeNear");
        if (ini.has("depthRangeFar")) farRange = ini.f("depthRangeFar");
        if (ini.has("gpuId")) gpuId = ini.i("gpuId");
        if (ini.has("orientationX") && ini.has("orientationZ")) {
            orientationX = ini.f("orientationX");
            orientationZ = ini.f("orientationZ");
        }
        refImageId = ini.i("refImageId");
        halfWindowSize = ini.i("halfWindowSize");
        numOfSamples = ini.i("numOfSamples");
        SPMAlpha = ini.f("SPMAlpha");
        gaussianSigma = ini.f("gaussianSigma");
        numOfIterations = ini.i("numOfIterations");
        outputFileName = ini.str("outputFileName");
        if (ini.has("blurMode")) {
            const std::string m = ini.str("blurMode");
            PatchMatch::setBlurMode(m == "none" ? PMX_BLUR_NONE : m == "box" ? PMX_BLUR_BOX : PMX_BLUR_GAUSS);
        }
        if (ini.has("initMode"))
            PatchMatch::setInitMode(ini.str("initMode") == "inverse" ? PMX_INIT_INVERSE : PMX_INIT_DEPTH);
        if (ini.has("rngSeed")) PatchMatch::setRngSeed((uint32_t)std::stoul(ini.str("rngSeed")));
    } catch (std::exception const& ex) {
        std::printf("Problems with config file: %s\n", ex.what());
        return 0;
    }

    setBestGPUDevice(gpuId);

    std::vector<Image> allImage;
    std::vector<std::pair<float, float>> depthRange;
    bool useNVM = false;
    if (readNVM(imageInfoFileName, allImage, depthRange)) {
        std::cout << "using NVM files" << std::endl;
        useNVM = true;
    } else if (readMiddleBurry(imageInfoFileName, allImage)) {
        std::cout << "using Middlebury files" << std::endl;
    } else {
        std::cout << "cannot parse image files" << std::endl;
        return EXIT_FAILURE;
    }
    const int blockDim_x = 32, blockDim_y = 16;
    if (useNVM) {
        if (refImageId < 0 || refImageId >= (int)depthRange.size()) {
            std::cout << "refImageId out of range" << std::endl;
            return EXIT_FAILURE;
        }
        nearRange = depthRange[refImageId].first;
        farRange = depthRange[refImageId].second;
    }

    PatchMatch pm(allImage, nearRange, farRange, halfWindowSize, blockDim_x, blockDim_y, refImageId,
                  numOfSamples, SPMAlpha, gaussianSigma, numOfIterations, orientationX, orientationZ);
    pm.runPatchMatch();
    pm._depthMap->saveToFile(outputFileName.c_str());
    if (ini.has("outputCostFileName")) pm._matchCost->saveToFile(ini.str("outputCostFileName"));
    if (ini.has("outputIdsFileName")) pm._usedImgsID->saveToFile(ini.str("outputIdsFileName"));
    return 0;
}
