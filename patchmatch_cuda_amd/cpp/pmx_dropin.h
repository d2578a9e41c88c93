/*
 * pmx_dropin.h -- the reference's host C++ classes (yhsa678/patchMatch_cuda,
 * src/patchMatch.h, src/image.h, src/array2d_wrapper.h, src/array2d_psng.h,
 * src/array2d_refImg.h, src/cudaArray_wrapper.h, src/cudaTranspose.h,
 * src/GaussianBlurCUDA.h, src/utility_CUDA.h) with the same names,
 * constructor signatures, methods and public data members, implemented on
 * the C ABI of libpmx.so (include/pmx.h).
 *
 * What changes for a caller:
 *   - CUDA types become HIP types (hipMemcpyKind, dim3 from HIP) and
 *     curandState becomes pmx_rng_state;
 *   - Image holds its camera as float arrays and its pixels in GrayImage
 *     (OpenCV is not used);
 *   - errors print "<message> in <file> at line <n>" and exit(EXIT_FAILURE),
 *     as CUDA_SAFE_CALL does (src/utility_CUDA.cpp:4-14).
 * Device memory of every wrapper is pitched device memory of the process-wide
 * engine context (pmx_dropin_context()).  Not re-entrant across threads, like
 * the reference.
 */
#ifndef PMX_DROPIN_H
#define PMX_DROPIN_H

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "pmx.h"

typedef unsigned char uchar;

/* ------------------------------------------------------------ utility */
/* The context every wrapper allocates on; created on first use on the
 * device chosen by setBestGPUDevice (device 0 if it was never called). */
pmx_ctx* pmx_dropin_context();
/* src/utility_CUDA.cpp:4-14 */
void _PMX_SAFE_CALL(int status, const char* file, int line);
#define PMX_SAFE_CALL(x) _PMX_SAFE_CALL((x), __FILE__, __LINE__)
/* src/utility_CUDA.cpp:73-112: gpuId >= 0 selects that device, otherwise
 * device 0 (a machine here carries one kind of GPU). */
void setBestGPUDevice(int gpuId);

/* ---------------------------------------------------------------- image */
/* Pixel buffer standing in for cv::Mat _imageData: 8-bit or float gray. */
struct GrayImage {
    int width = 0;
    int height = 0;
    int format = PMX_IMG_U8;           /* PMX_IMG_U8 or PMX_IMG_F32 */
    std::vector<unsigned char> u8;     /* width*height when format == U8 */
    std::vector<float> f32;            /* width*height when format == F32 */
};

/* src/image.h:10-33 */
class Image {
public:
    float _R[9];
    float _C[3];        /* C = -R^T T */
    float _T[3];        /* T = -R C */
    float _K[9];
    float _proj[12];    /* K [R | T], row-major 3x4 */
    float _inverseK[9];
    float _inverseR[9];
    float _relative_R[9];
    float _relative_T[3];
    float _H1[9];
    float _H2[9];
    GrayImage _imageData;

    /* src/image.cpp:16-38 / :40-62 (image given in memory) */
    void updateCamParam(float* K, float* R, float* T, const GrayImage& image);
    /* src/image.cpp:4-13, composed in FP64 and stored as float */
    void init_relative(const Image& refImg, float orientationX, float orientationZ);
    /* The camera as the C ABI takes it (FP64 K, R and centre C). */
    pmx_camera camera() const;
};

/* ------------------------------------------------------- device arrays */
/* src/array2d_wrapper.h:14-121: width x (height*depth) pitched device array. */
template <class T>
class Array2D_wrapper {
public:
    T* _array2D;
    int _pitchData;

    Array2D_wrapper(int width, int height, int blockDim_x, int blockDim_y, int depth = 1);
    virtual ~Array2D_wrapper();

    /* src/array2d_wrapper.h:137-145 (T = float only) */
    void randNumGen(float rangeStart, float rangeEnd, pmx_rng_state* devStates, int pitchState);
    int getWidth();
    int getHeight();
    int getDepth();
    /* src/array2d_wrapper.h:41-49 */
    void copyData(T* ptr, int pitch, hipMemcpyKind kind);
    /* src/array2d_wrapper.h:51-92: "W&H&D&" then the raw elements */
    void saveToFile(std::string fileName);
    /* src/array2d_wrapper.h:94-103: one layer as text, one value a line */
    void saveToFile(std::string fileName, int layer);

protected:
    /* a non-owning view of engine memory (PatchMatch's outputs) */
    Array2D_wrapper(T* dev, int pitch, int width, int height, int depth, int blockDim_x,
                    int blockDim_y);
    void computeCUDAConfig();

    int _height;
    int _depth;
    int _width;
    dim3 _blockSize;
    int _blockDim_x;
    int _blockDim_y;
    dim3 _gridSize;
    bool _owner;

    friend class PatchMatch;
};

/* src/array2d_psng.h:8-19: one generator state per pixel, seeded at
 * construction (block-linear id, the project xorwow generator). */
class Array2D_psng : public Array2D_wrapper<pmx_rng_state> {
    void randStateGen();

public:
    Array2D_psng(int width, int height, int blockDim_x, int blockDim_y, uint32_t seed = 7u);

private:
    /* a non-owning view of an engine's generator states (PatchMatch::_psngState) */
    Array2D_psng(pmx_rng_state* dev, int pitch, int width, int height, int blockDim_x,
                 int blockDim_y);
    uint32_t _seed;
    friend class PatchMatch;
};

/* src/cudaArray_wrapper.h:6-25: a layered width x height x depth image.
 * _array3D is a packed device block (layer k at k*width*height elements of
 * the copied type) instead of a texture array. */
class CudaArray_wrapper {
    int _width;
    int _height;
    int _depth;

public:
    void* _array3D;

    CudaArray_wrapper(int width, int height, int depth);
    ~CudaArray_wrapper();

    template <typename T>
    void array3DCopy(T* img, hipMemcpyKind kind);
    void array3DCopy_float(float* data, hipMemcpyKind kind, int dataPitch);

private:
    size_t _elem;
    void alloc(size_t elem);
};

/* src/array2d_refImg.h:7-62 */
class Array2d_refImg {
public:
    Array2D_wrapper<float>* _refImage_sum_II;
    Array2D_wrapper<float>* _refImage_sum_I;
    Array2D_wrapper<float>* _refImageData;

    int _refWidth;
    int _refHeight;

    CudaArray_wrapper* _tempArray;
    CudaArray_wrapper* _tempArray_float;

    Array2d_refImg(int refWidth, int refHeight, int blockDim_x, int blockDim_y, unsigned char* img);
    Array2d_refImg(int refWidth, int refHeight, int blockDim_x, int blockDim_y);
    ~Array2d_refImg();

    /* src/array2d_refImg.cu:70-131: I = u/255 (or the float image), window
     * sums of I and I^2 with zero border; hw outside 1..22 prints the
     * reference's warning and exits with status 0 (:96). */
    void filterImage(int halfWindowSize);

private:
    /* views of an engine's reference image and window sums
     * (PatchMatch::_refImage); takes the three wrapper objects */
    Array2d_refImg(Array2D_wrapper<float>* I, Array2D_wrapper<float>* sum_I,
                   Array2D_wrapper<float>* sum_II, int refWidth, int refHeight);
    friend class PatchMatch;
};

/* src/cudaTranspose.h:52-71 (4-byte element types) */
class cudaTranspose {
public:
    template <class T>
    static void transpose2dData(T* input, T* output, int width, int height, int pitchInput,
                                int pitchOutput);
};

/* src/GaussianBlurCUDA.h:5-24.  The reference's CreateFilterKernel replaces
 * the Gaussian taps by uniform 1/width ones (src/GaussianBlurCUDA.cu:120-129);
 * useBoxTaps keeps that behaviour, the default is the normalised Gaussian. */
class GaussianBlurCUDA {
public:
    GaussianBlurCUDA(int width, int height, float sigma, bool useBoxTaps = false);
    ~GaussianBlurCUDA();
    /* host src -> host dst, one width x height image */
    int Filter(float* dst, const float* src);
    /* in place on `depth` stacked pitched device planes */
    int FilterMultipleImages(float* data, int pitch, int depth);

private:
    int m_nWidth, m_nHeight, m_nKernelWidth;
    float m_paraSigma;
    int m_mode;
};

/* --------------------------------------------------------- PatchMatch */
/* src/patchMatch.h:13-87.  The constructor validates, uploads the images
 * and runs the constructor-time device work; runPatchMatch runs the
 * iterations.  _depthMap (W x H), _matchCost (W x H*V, view planes stacked)
 * and _usedImgsID (W x H*S, uchar) are views of the engine's outputs in the
 * reference layouts, valid after runPatchMatch until destruction. */
class PatchMatch {
public:
    float _nearRange;
    float _farRange;
    int _halfWindowSize;
    int _blockDim_x;
    int _blockDim_y;
    int _refImageId;
    dim3 _blockSize;
    dim3 _gridSize;

    Array2D_wrapper<float>* _depthMap;
    Array2D_wrapper<float>* _matchCost;
    Array2D_wrapper<uchar>* _usedImgsID;
    /* src/patchMatch.h:41, :49: the reference image (pre-blurred) with its
     * window sums, and the per-pixel generator states; views of the
     * engine's device memory, valid until destruction */
    Array2d_refImg* _refImage;
    Array2D_psng* _psngState;

    int _numOfTargetImages;
    int _maxWidth;
    int _maxHeight;
    int _refWidth;
    int _refHeight;
    float* _transformHH;    /* V x 18 (H1, H2), host, as src/patchMatch.cu:118-129 */
    int _numOfSamples;
    float _SPMAlpha;
    float _gaussianSigma;
    int _numOfIterations;
    float _orientationX;
    float _orientationZ;

    PatchMatch(std::vector<Image>& allImage, float nearRange, float farRange, int halfWindowSize,
               int blockDim_x, int blockDim_y, int refImageId, int numOfSamples, float SPMAlpha,
               float gaussianSigma, int numOfIterations, float orientationX, float orientationZ);
    ~PatchMatch();

    void runPatchMatch();

    /* Engine settings with no reference counterpart. */
    static void setBlurMode(int pmx_blur_mode);    /* default PMX_BLUR_GAUSS */
    static void setInitMode(int pmx_init_mode);    /* default PMX_INIT_DEPTH (reference) */
    static void setRngSeed(uint32_t seed);         /* default 7 */
    pmx_engine* engine() { return _engine; }

private:
    pmx_engine* _engine;
    void computeCUDAConfig(int width, int height, int blockDim_x, int blockDim_y);
    void bindOutputs();
};

#endif
