/*
 * dropin_helpers.cpp -- drives the reference's helper classes through
 * libpmx_drop_in.so (the classes of src/array2d_psng.h, src/array2d_wrapper.h,
 * src/array2d_refImg.h, src/cudaTranspose.h, src/GaussianBlurCUDA.h,
 * src/cudaArray_wrapper.h and the PatchMatch members _refImage/_psngState of
 * src/patchMatch.h:41-49) and writes every result to a binary file, which
 * tests/test_dropin_helpers_gpu.py compares with the FP64 oracle.
 *
 *   dropin_helpers <input> <output>
 *
 * Input (little endian): int32 W, H, hw, block_x, block_y; uint32 seed;
 * float near, far, sigma; uint8 image W*H; float planes 2*W*H; int32 Wt, Ht;
 * float Wt*Ht; int32 n_images, S, iterations; float alpha; then per image
 * float K[9], R[9], T[3] and float pixels W*H (image 0 is the reference).
 * Output: the arrays in the order written below, raw, without headers.
 */
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "pmx_dropin.h"

namespace {

struct Reader {
    std::ifstream in;
    explicit Reader(const char* path) : in(path, std::ios::binary) {}
    template <class T>
    T get() {
        T v;
        in.read(reinterpret_cast<char*>(&v), sizeof(T));
        return v;
    }
    template <class T>
    std::vector<T> vec(size_t n) {
        std::vector<T> v(n);
        in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
        return v;
    }
};

template <class T>
void put(std::ofstream& out, const std::vector<T>& v) {
    out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

template <class T>
std::vector<T> download(Array2D_wrapper<T>& a, int width, int rows) {
    std::vector<T> h((size_t)width * rows);
    a.copyData(h.data(), width * (int)sizeof(T), hipMemcpyDeviceToHost);
    return h;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: dropin_helpers <input> <output>" << std::endl;
        return 2;
    }
    Reader r(argv[1]);
    const int W = r.get<int>(), H = r.get<int>(), hw = r.get<int>();
    const int bx = r.get<int>(), by = r.get<int>();
    const uint32_t seed = r.get<uint32_t>();
    const float near_d = r.get<float>(), far_d = r.get<float>(), sigma = r.get<float>();
    std::vector<uint8_t> img8 = r.vec<uint8_t>((size_t)W * H);
    std::vector<float> planes = r.vec<float>((size_t)2 * W * H);
    const int Wt = r.get<int>(), Ht = r.get<int>();
    std::vector<float> tin = r.vec<float>((size_t)Wt * Ht);
    const int n_img = r.get<int>(), S = r.get<int>(), iters = r.get<int>();
    const float alpha = r.get<float>();
    std::vector<Image> images(n_img);
    for (int i = 0; i < n_img; i++) {
        std::vector<float> K = r.vec<float>(9), R = r.vec<float>(9), T = r.vec<float>(3);
        GrayImage g;
        g.width = W;
        g.height = H;
        g.format = PMX_IMG_F32;
        g.f32 = r.vec<float>((size_t)W * H);
        images[i].updateCamParam(K.data(), R.data(), T.data(), g);
    }
    if (!r.in) {
        std::cerr << "short input" << std::endl;
        return 2;
    }
    std::ofstream out(argv[2], std::ios::binary);

    /* 1. Array2D_psng: setup_kernel states (src/array2d_psng.cu:3-31) */
    Array2D_psng psng(W, H, bx, by, seed);
    put(out, download(psng, W, H));
    /* 2. randNumGen (src/array2d_wrapper.h:137-145): depth and states after */
    Array2D_wrapper<float> depth(W, H, bx, by);
    depth.randNumGen(near_d, far_d, psng._array2D, psng._pitchData);
    put(out, download(depth, W, H));
    put(out, download(psng, W, H));
    /* 3. Array2d_refImg::filterImage (src/array2d_refImg.cu:70-131) */
    Array2d_refImg ref(W, H, bx, by, img8.data());
    ref.filterImage(hw);
    put(out, download(*ref._refImageData, W, H));
    put(out, download(*ref._refImage_sum_I, W, H));
    put(out, download(*ref._refImage_sum_II, W, H));
    /* 4. cudaTranspose::transpose2dData (src/cudaTranspose.h:52-71) */
    Array2D_wrapper<float> tsrc(Wt, Ht, bx, by), tdst(Ht, Wt, bx, by);
    tsrc.copyData(tin.data(), Wt * (int)sizeof(float), hipMemcpyHostToDevice);
    cudaTranspose::transpose2dData<float>(tsrc._array2D, tdst._array2D, Wt, Ht, tsrc._pitchData,
                                          tdst._pitchData);
    put(out, download(tdst, Ht, Wt));
    /* 5. GaussianBlurCUDA::FilterMultipleImages (src/GaussianBlurCUDA.cu:189-199),
     *    normalised Gaussian taps and the reference's uniform taps */
    for (int box = 0; box < 2; box++) {
        Array2D_wrapper<float> stack(W, H, bx, by, 2);
        stack.copyData(planes.data(), W * (int)sizeof(float), hipMemcpyHostToDevice);
        GaussianBlurCUDA blur(W, H, sigma, box != 0);
        blur.FilterMultipleImages(stack._array2D, stack._pitchData, 2);
        put(out, download(stack, W, 2 * H));
    }
    /* 6. CudaArray_wrapper::array3DCopy (src/cudaArray_wrapper.h:27-60) round trips */
    {
        CudaArray_wrapper ca(W, H, 2);
        ca.array3DCopy<float>(planes.data(), hipMemcpyHostToDevice);
        std::vector<float> back((size_t)2 * W * H, -1.0f);
        ca.array3DCopy<float>(back.data(), hipMemcpyDeviceToHost);
        put(out, back);
        CudaArray_wrapper cu(W, H, 1);
        cu.array3DCopy<unsigned char>(img8.data(), hipMemcpyHostToDevice);
        std::vector<uint8_t> back8((size_t)W * H, 0);
        cu.array3DCopy<unsigned char>(back8.data(), hipMemcpyDeviceToHost);
        put(out, back8);
    }
    /* 7. PatchMatch::_refImage and _psngState after the constructor
     *    (src/patchMatch.cu:132-206), then the depth map of a run */
    {
        PatchMatch pm(images, near_d, far_d, hw, bx, by, 0, S, alpha, sigma, iters, 0.0f, 1.0f);
        put(out, download(*pm._refImage->_refImageData, W, H));
        put(out, download(*pm._refImage->_refImage_sum_I, W, H));
        put(out, download(*pm._refImage->_refImage_sum_II, W, H));
        put(out, download(*pm._psngState, W, H));
        pm.runPatchMatch();
        put(out, download(*pm._depthMap, W, H));
        put(out, download(*pm._psngState, W, H));
    }
    out.close();
    std::cout << "dropin_helpers: done" << std::endl;
    return 0;
}
