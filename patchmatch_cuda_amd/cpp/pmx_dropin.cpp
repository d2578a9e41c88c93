/*
 * pmx_dropin.cpp -- reference-compatible host classes on the C ABI of
 * libpmx.so.  Each method cites the reference member it stands in for.
 */
#include "pmx_dropin.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <type_traits>

namespace {

pmx_ctx* g_ctx = nullptr;
int g_device = 0;
int g_blur_mode = PMX_BLUR_GAUSS;
int g_init_mode = PMX_INIT_DEPTH;
uint32_t g_seed = 7u;

int memcpy_kind(hipMemcpyKind kind) {
    switch (kind) {
    case hipMemcpyHostToDevice: return PMX_MEMCPY_H2D;
    case hipMemcpyDeviceToHost: return PMX_MEMCPY_D2H;
    case hipMemcpyDeviceToDevice: return PMX_MEMCPY_D2D;
    default: return 0;
    }
}

/* 3x3 helpers in FP64 (row-major) */
void mul3(const double* a, const double* b, double* o) {
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            t[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
    std::memcpy(o, t, sizeof(t));
}

bool inv3(const double* m, double* o) {
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8],
                 c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (det == 0.0) return false;
    const double r = 1.0 / det;
    o[0] = c00 * r;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * r;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * r;
    o[3] = c01 * r;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * r;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * r;
    o[6] = c02 * r;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * r;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * r;
    return true;
}

template <int N>
void to_double(const float* f, double* d) {
    for (int i = 0; i < N; i++) d[i] = (double)f[i];
}

template <int N>
void to_float(const double* d, float* f) {
    for (int i = 0; i < N; i++) f[i] = (float)d[i];
}

/* float intensities of a GrayImage: u/255 for 8-bit images, as a texture
 * read with cudaReadModeNormalizedFloat */
std::vector<float> as_float(const GrayImage& g) {
    if (g.format == PMX_IMG_F32) return g.f32;
    std::vector<float> out(g.u8.size());
    for (size_t i = 0; i < out.size(); i++) out[i] = (float)g.u8[i] / 255.0f;
    return out;
}

} // namespace

/* ------------------------------------------------------------ utility */

void _PMX_SAFE_CALL(int status, const char* file, int line) {
    if (status != PMX_OK) {
        std::printf("%s in %s at line %i\n", pmx_last_error(g_ctx), file, line);
        std::fflush(stdout);
        std::exit(EXIT_FAILURE);
    }
}

pmx_ctx* pmx_dropin_context() {
    if (!g_ctx) _PMX_SAFE_CALL(pmx_ctx_create(g_device, &g_ctx), __FILE__, __LINE__);
    return g_ctx;
}

void setBestGPUDevice(int gpuId) {
    int n = 0;
    pmx_device_count(&n);
    if (n == 0) {
        std::printf("No GPU is detected in the machine");
        std::exit(EXIT_FAILURE);
    }
    const int dev = (gpuId >= 0 && gpuId < n) ? gpuId : 0;
    if (g_ctx && dev != g_device) {
        pmx_ctx_destroy(g_ctx);
        g_ctx = nullptr;
    }
    g_device = dev;
    std::cout << "the device id is: " << dev << std::endl;
    pmx_dropin_context();
}

/* ---------------------------------------------------------------- Image */

void Image::updateCamParam(float* K, float* R, float* T, const GrayImage& image) {
    std::memcpy(_K, K, sizeof(_K));
    std::memcpy(_R, R, sizeof(_R));
    std::memcpy(_T, T, sizeof(_T));
    double Kd[9], Rd[9], Td[3], Ki[9];
    to_double<9>(_K, Kd);
    to_double<9>(_R, Rd);
    to_double<3>(_T, Td);
    for (int i = 0; i < 3; i++) /* C = -R^T T */
        _C[i] = (float)(-(Rd[i] * Td[0] + Rd[3 + i] * Td[1] + Rd[6 + i] * Td[2]));
    for (int i = 0; i < 3; i++) { /* proj = K [R | T] */
        for (int j = 0; j < 4; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += Kd[i * 3 + k] * (j < 3 ? Rd[k * 3 + j] : Td[k]);
            _proj[i * 4 + j] = (float)s;
        }
    }
    if (!inv3(Kd, Ki)) {
        std::printf("singular camera matrix in %s at line %i\n", __FILE__, __LINE__);
        std::exit(EXIT_FAILURE);
    }
    to_float<9>(Ki, _inverseK);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) _inverseR[i * 3 + j] = _R[j * 3 + i];
    _imageData = image;
}

void Image::init_relative(const Image& refImg, float orientationX, float orientationZ) {
    double R[9], Rri[9], K[9], Kri[9], C[3], Cr[3], rel[9], t[9];
    to_double<9>(_R, R);
    to_double<9>(refImg._inverseR, Rri);
    to_double<9>(_K, K);
    to_double<9>(refImg._inverseK, Kri);
    to_double<3>(_C, C);
    to_double<3>(refImg._C, Cr);
    mul3(R, Rri, rel);
    double T[3];
    for (int i = 0; i < 3; i++)
        T[i] = R[i * 3] * (C[0] - Cr[0]) + R[i * 3 + 1] * (C[1] - Cr[1]) + R[i * 3 + 2] * (C[2] - Cr[2]);
    to_float<9>(rel, _relative_R);
    to_float<3>(T, _relative_T);
    mul3(K, rel, t);
    double H1[9], H2[9], tn[9];
    mul3(t, Kri, H1);
    const double n[3] = {(double)orientationX, 0.0, (double)orientationZ};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tn[i * 3 + j] = T[i] * n[j];
    mul3(K, tn, t);
    mul3(t, Kri, H2);
    to_float<9>(H1, _H1);
    to_float<9>(H2, _H2);
}

pmx_camera Image::camera() const {
    pmx_camera c;
    to_double<9>(_K, c.K);
    to_double<9>(_R, c.R);
    to_double<3>(_C, c.C);
    return c;
}

/* -------------------------------------------------------- Array2D_wrapper */

template <class T>
Array2D_wrapper<T>::Array2D_wrapper(int width, int height, int blockDim_x, int blockDim_y, int depth)
    : _array2D(nullptr), _pitchData(0), _height(height), _depth(depth), _width(width),
      _blockDim_x(blockDim_x), _blockDim_y(blockDim_y), _owner(true) {
    void* p = nullptr;
    size_t pitch = 0;
    PMX_SAFE_CALL(pmx_malloc_pitch(pmx_dropin_context(), &p, &pitch, (size_t)_width * sizeof(T),
                                   (size_t)_height * (size_t)_depth));
    _array2D = static_cast<T*>(p);
    _pitchData = static_cast<int>(pitch);
    computeCUDAConfig();
}

template <class T>
Array2D_wrapper<T>::Array2D_wrapper(T* dev, int pitch, int width, int height, int depth,
                                    int blockDim_x, int blockDim_y)
    : _array2D(dev), _pitchData(pitch), _height(height), _depth(depth), _width(width),
      _blockDim_x(blockDim_x), _blockDim_y(blockDim_y), _owner(false) {
    computeCUDAConfig();
}

template <class T>
Array2D_wrapper<T>::~Array2D_wrapper() {
    if (_owner && _array2D && g_ctx) pmx_free(g_ctx, _array2D);
}

template <class T>
int Array2D_wrapper<T>::getWidth() { return _width; }
template <class T>
int Array2D_wrapper<T>::getHeight() { return _height; }
template <class T>
int Array2D_wrapper<T>::getDepth() { return _depth; }

template <class T>
void Array2D_wrapper<T>::computeCUDAConfig() {
    _blockSize = dim3(_blockDim_x, _blockDim_y, 1);
    _gridSize = dim3((_width - 1) / _blockDim_x + 1, (_height - 1) / _blockDim_y + 1, 1);
}

template <class T>
void Array2D_wrapper<T>::copyData(T* ptr, int pitch, hipMemcpyKind kind) {
    const size_t wb = (size_t)_width * sizeof(T), rows = (size_t)_height * (size_t)_depth;
    if (kind == hipMemcpyDeviceToHost)
        PMX_SAFE_CALL(pmx_memcpy2d(pmx_dropin_context(), ptr, (size_t)pitch, _array2D,
                                   (size_t)_pitchData, wb, rows, PMX_MEMCPY_D2H));
    else
        PMX_SAFE_CALL(pmx_memcpy2d(pmx_dropin_context(), _array2D, (size_t)_pitchData, ptr,
                                   (size_t)pitch, wb, rows, memcpy_kind(kind)));
}

template <class T>
void Array2D_wrapper<T>::saveToFile(std::string fileName) {
    {
        std::ofstream fout(fileName, std::ios_base::out | std::ios_base::trunc);
        fout << _width << "&" << _height << "&" << _depth << "&";
    }
    std::vector<T> host((size_t)_width * _height * _depth);
    copyData(host.data(), _width * (int)sizeof(T), hipMemcpyDeviceToHost);
    std::ofstream fb(fileName, std::ios_base::out | std::ios_base::binary | std::ios_base::app);
    fb.write(reinterpret_cast<const char*>(host.data()), (std::streamsize)(sizeof(T) * host.size()));
}

template <class T>
void Array2D_wrapper<T>::saveToFile(std::string fileName, int layer) {
    std::vector<T> host((size_t)_width * _height);
    PMX_SAFE_CALL(pmx_memcpy2d(pmx_dropin_context(), host.data(), (size_t)_width * sizeof(T),
                               reinterpret_cast<const char*>(_array2D) +
                                   (size_t)layer * _height * _pitchData,
                               (size_t)_pitchData, (size_t)_width * sizeof(T), (size_t)_height,
                               PMX_MEMCPY_D2H));
    std::ofstream fout(fileName, std::ios_base::out | std::ios_base::trunc);
    if constexpr (std::is_arithmetic<T>::value) {
        std::copy(host.begin(), host.end(), std::ostream_iterator<T>(fout, "\n"));
    } else { /* generator states: the six words of each, one state a line */
        for (const T& s : host) {
            const uint32_t* w = reinterpret_cast<const uint32_t*>(&s);
            for (size_t k = 0; k < sizeof(T) / 4; k++) fout << w[k] << (k + 1 < sizeof(T) / 4 ? " " : "\n");
        }
    }
}

template <class T>
void Array2D_wrapper<T>::randNumGen(float rangeStart, float rangeEnd, pmx_rng_state* devStates,
                                    int pitchState) {
    /* generate_kernel_float(_withDepth): one draw per pixel and layer, in
     * layer order, from the pixel's state */
    if (!std::is_same<T, float>::value) {
        std::printf("randNumGen fills float arrays only, in %s at line %i\n", __FILE__, __LINE__);
        std::exit(EXIT_FAILURE);
    }
    for (int k = 0; k < _depth; k++) {
        float* layer = reinterpret_cast<float*>(reinterpret_cast<char*>(_array2D) +
                                                (size_t)k * _height * _pitchData);
        PMX_SAFE_CALL(pmx_rng_fill_depth(pmx_dropin_context(), devStates, pitchState, layer,
                                         _pitchData, _width, _height, rangeStart, rangeEnd,
                                         PMX_INIT_DEPTH));
    }
}

template class Array2D_wrapper<float>;
template class Array2D_wrapper<uchar>;
template class Array2D_wrapper<pmx_rng_state>;

/* ---------------------------------------------------------- Array2D_psng */

Array2D_psng::Array2D_psng(int width, int height, int blockDim_x, int blockDim_y, uint32_t seed)
    : Array2D_wrapper<pmx_rng_state>(width, height, blockDim_x, blockDim_y), _seed(seed) {
    randStateGen();
}

Array2D_psng::Array2D_psng(pmx_rng_state* dev, int pitch, int width, int height, int blockDim_x,
                           int blockDim_y)
    : Array2D_wrapper<pmx_rng_state>(dev, pitch, width, height, 1, blockDim_x, blockDim_y),
      _seed(0) {}

void Array2D_psng::randStateGen() {
    /* src/array2d_psng.cu:20-31 */
    PMX_SAFE_CALL(pmx_rng_init(pmx_dropin_context(), _array2D, _pitchData, _width, _height,
                               _blockDim_x, _blockDim_y, _seed));
}

/* ------------------------------------------------------ CudaArray_wrapper */

CudaArray_wrapper::CudaArray_wrapper(int width, int height, int depth)
    : _width(width), _height(height), _depth(depth), _array3D(nullptr), _elem(0) {}

CudaArray_wrapper::~CudaArray_wrapper() {
    if (_array3D && g_ctx) pmx_free(g_ctx, _array3D);
}

void CudaArray_wrapper::alloc(size_t elem) {
    if (_array3D && elem == _elem) return;
    if (_array3D) pmx_free(pmx_dropin_context(), _array3D);
    _elem = elem;
    PMX_SAFE_CALL(pmx_malloc(pmx_dropin_context(), &_array3D,
                             elem * (size_t)_width * (size_t)_height * (size_t)_depth));
}

template <typename T>
void CudaArray_wrapper::array3DCopy(T* img, hipMemcpyKind kind) {
    /* src/cudaArray_wrapper.h:27-60: tightly packed host layers */
    const size_t wb = (size_t)_width * sizeof(T), rows = (size_t)_height * (size_t)_depth;
    if (kind == hipMemcpyHostToDevice) {
        alloc(sizeof(T));
        PMX_SAFE_CALL(pmx_memcpy2d(pmx_dropin_context(), _array3D, wb, img, wb, wb, rows,
                                   PMX_MEMCPY_H2D));
    } else if (kind == hipMemcpyDeviceToHost && _array3D) {
        PMX_SAFE_CALL(pmx_memcpy2d(pmx_dropin_context(), img, wb, _array3D, wb, wb, rows,
                                   PMX_MEMCPY_D2H));
    }
}

template void CudaArray_wrapper::array3DCopy<unsigned char>(unsigned char*, hipMemcpyKind);
template void CudaArray_wrapper::array3DCopy<float>(float*, hipMemcpyKind);

void CudaArray_wrapper::array3DCopy_float(float* data, hipMemcpyKind kind, int dataPitch) {
    /* src/cudaArray_wrapper.cpp:13-41 */
    alloc(sizeof(float));
    const size_t wb = (size_t)_width * sizeof(float), rows = (size_t)_height * (size_t)_depth;
    if (kind == hipMemcpyDeviceToDevice)
        PMX_SAFE_CALL(pmx_memcpy2d(pmx_dropin_context(), _array3D, wb, data, (size_t)dataPitch, wb,
                                   rows, PMX_MEMCPY_D2D));
    else if (kind == hipMemcpyHostToDevice)
        PMX_SAFE_CALL(pmx_memcpy2d(pmx_dropin_context(), _array3D, wb, data, wb, wb, rows,
                                   PMX_MEMCPY_H2D));
}

/* --------------------------------------------------------- Array2d_refImg */

Array2d_refImg::Array2d_refImg(int refWidth, int refHeight, int blockDim_x, int blockDim_y,
                               unsigned char* img)
    : Array2d_refImg(refWidth, refHeight, blockDim_x, blockDim_y) {
    _tempArray->array3DCopy<unsigned char>(img, hipMemcpyHostToDevice);
}

Array2d_refImg::Array2d_refImg(int refWidth, int refHeight, int blockDim_x, int blockDim_y)
    : _refImage_sum_II(nullptr), _refImage_sum_I(nullptr), _refImageData(nullptr),
      _refWidth(refWidth), _refHeight(refHeight), _tempArray(nullptr), _tempArray_float(nullptr) {
    _refImage_sum_II = new Array2D_wrapper<float>(refWidth, refHeight, blockDim_x, blockDim_y);
    _refImage_sum_I = new Array2D_wrapper<float>(refWidth, refHeight, blockDim_x, blockDim_y);
    _refImageData = new Array2D_wrapper<float>(refWidth, refHeight, blockDim_x, blockDim_y);
    _tempArray = new CudaArray_wrapper(refWidth, refHeight, 1);
    _tempArray_float = new CudaArray_wrapper(refWidth, refHeight, 1);
}

Array2d_refImg::Array2d_refImg(Array2D_wrapper<float>* I, Array2D_wrapper<float>* sum_I,
                               Array2D_wrapper<float>* sum_II, int refWidth, int refHeight)
    : _refImage_sum_II(sum_II), _refImage_sum_I(sum_I), _refImageData(I), _refWidth(refWidth),
      _refHeight(refHeight), _tempArray(nullptr), _tempArray_float(nullptr) {}

Array2d_refImg::~Array2d_refImg() {
    delete _refImage_sum_I;
    delete _refImage_sum_II;
    delete _refImageData;
    delete _tempArray;
    delete _tempArray_float;
}

void Array2d_refImg::filterImage(int halfWindowSize) {
    if (halfWindowSize < 1 || halfWindowSize > 22) {
        /* src/array2d_refImg.cu:96 */
        std::cout << "Warning: too large size of windows" << std::endl;
        std::exit(0);
    }
    /* intensities: the uchar image read as u/255, or a float image placed in
     * _tempArray_float by the caller */
    std::vector<float> I((size_t)_refWidth * _refHeight);
    if (_tempArray->_array3D) {
        std::vector<unsigned char> u(I.size());
        _tempArray->array3DCopy<unsigned char>(u.data(), hipMemcpyDeviceToHost);
        for (size_t i = 0; i < I.size(); i++) I[i] = (float)u[i] / 255.0f;
    } else if (_tempArray_float->_array3D) {
        _tempArray_float->array3DCopy<float>(I.data(), hipMemcpyDeviceToHost);
    } else {
        std::printf("reference image not loaded in %s at line %i\n", __FILE__, __LINE__);
        std::exit(EXIT_FAILURE);
    }
    Array2D_wrapper<float> src(_refWidth, _refHeight, 32, 16);
    src.copyData(I.data(), _refWidth * (int)sizeof(float), hipMemcpyHostToDevice);
    PMX_SAFE_CALL(pmx_box_sums(pmx_dropin_context(), src._array2D, src._pitchData, _refWidth,
                               _refHeight, halfWindowSize, _refImageData->_array2D,
                               _refImage_sum_I->_array2D, _refImage_sum_II->_array2D,
                               _refImageData->_pitchData));
}

/* ----------------------------------------------------------- cudaTranspose */

template <class T>
void cudaTranspose::transpose2dData(T* input, T* output, int width, int height, int pitchInput,
                                    int pitchOutput) {
    static_assert(sizeof(T) == 4, "transpose2dData moves 4-byte elements");
    PMX_SAFE_CALL(pmx_transpose_f32(pmx_dropin_context(), reinterpret_cast<const float*>(input),
                                    reinterpret_cast<float*>(output), width, height, pitchInput,
                                    pitchOutput));
}

template void cudaTranspose::transpose2dData<float>(float*, float*, int, int, int, int);
template void cudaTranspose::transpose2dData<int>(int*, int*, int, int, int, int);

/* -------------------------------------------------------- GaussianBlurCUDA */

GaussianBlurCUDA::GaussianBlurCUDA(int width, int height, float sigma, bool useBoxTaps)
    : m_nWidth(width), m_nHeight(height), m_paraSigma(sigma),
      m_mode(useBoxTaps ? PMX_BLUR_BOX : PMX_BLUR_GAUSS) {
    /* src/GaussianBlurCUDA.cu:102-115 */
    int w = (int)(5.0f * sigma);
    if (w % 2 == 0) w += 1;
    m_nKernelWidth = std::min(45, std::max(5, w));
}

GaussianBlurCUDA::~GaussianBlurCUDA() {}

int GaussianBlurCUDA::Filter(float* dst, const float* src) {
    Array2D_wrapper<float> buf(m_nWidth, m_nHeight, 32, 16);
    buf.copyData(const_cast<float*>(src), m_nWidth * (int)sizeof(float), hipMemcpyHostToDevice);
    PMX_SAFE_CALL(pmx_blur(pmx_dropin_context(), buf._array2D, buf._pitchData, m_nWidth, m_nHeight,
                           1, m_paraSigma, m_mode));
    buf.copyData(dst, m_nWidth * (int)sizeof(float), hipMemcpyDeviceToHost);
    return 0;
}

int GaussianBlurCUDA::FilterMultipleImages(float* data, int pitch, int depth) {
    PMX_SAFE_CALL(pmx_blur(pmx_dropin_context(), data, pitch, m_nWidth, m_nHeight, depth,
                           m_paraSigma, m_mode));
    return 0;
}

/* -------------------------------------------------------------- PatchMatch */

void PatchMatch::setBlurMode(int mode) { g_blur_mode = mode; }
void PatchMatch::setInitMode(int mode) { g_init_mode = mode; }
void PatchMatch::setRngSeed(uint32_t seed) { g_seed = seed; }

PatchMatch::PatchMatch(std::vector<Image>& allImage, float nearRange, float farRange,
                       int halfWindowSize, int blockDim_x, int blockDim_y, int refImageId,
                       int numOfSamples, float SPMAlpha, float gaussianSigma, int numOfIterations,
                       float orientationX, float orientationZ)
    : _nearRange(nearRange), _farRange(farRange), _halfWindowSize(halfWindowSize),
      _blockDim_x(blockDim_x), _blockDim_y(blockDim_y), _refImageId(refImageId),
      _depthMap(nullptr), _matchCost(nullptr), _usedImgsID(nullptr), _refImage(nullptr),
      _psngState(nullptr), _maxWidth(0), _maxHeight(0),
      _transformHH(nullptr), _numOfSamples(numOfSamples), _SPMAlpha(SPMAlpha),
      _gaussianSigma(gaussianSigma), _numOfIterations(numOfIterations),
      _orientationX(orientationX), _orientationZ(orientationZ), _engine(nullptr) {
    /* src/patchMatch.cu:137-142 */
    _numOfTargetImages = static_cast<int>(allImage.size()) - 1;
    if (_numOfTargetImages <= 0) {
        std::cout << "Error: at least 2 images are needed for stereo" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    if (refImageId < 0 || refImageId > _numOfTargetImages) {
        std::printf("reference image id %d out of range in %s at line %i\n", refImageId, __FILE__,
                    __LINE__);
        std::exit(EXIT_FAILURE);
    }
    /* the constructor's filterImage accepts 1..22 (src/array2d_refImg.cu:96) */
    if (halfWindowSize < 1 || halfWindowSize > 22) {
        std::cout << "Warning: too large size of windows" << std::endl;
        std::exit(0);
    }
    /* src/patchMatch.cu:143-145 */
    for (size_t i = 0; i < allImage.size(); i++)
        allImage[i].init_relative(allImage[refImageId], _orientationX, _orientationZ);

    pmx_params p{};
    p.half_window = halfWindowSize;
    p.num_samples = numOfSamples;
    p.num_iterations = numOfIterations;
    p.depth_near = nearRange;
    p.depth_far = farRange;
    p.spm_alpha = SPMAlpha;
    p.gaussian_sigma = gaussianSigma;
    p.blur_mode = g_blur_mode;
    p.init_mode = g_init_mode;
    p.orientation_x = orientationX;
    p.orientation_z = orientationZ;
    p.block_x = blockDim_x;
    p.block_y = blockDim_y;
    p.rng_seed = g_seed;

    /* copyData (src/patchMatch.cu:51-130): the reference plus the others in
     * order; all images go up as u8 when every image is u8, else as float */
    bool all_u8 = true;
    for (const Image& im : allImage) all_u8 = all_u8 && im._imageData.format == PMX_IMG_U8;
    const int fmt = all_u8 ? PMX_IMG_U8 : PMX_IMG_F32;
    std::vector<std::vector<float>> fbuf;
    auto pixels = [&](const Image& im) -> const void* {
        if (fmt == PMX_IMG_U8) return im._imageData.u8.data();
        if (im._imageData.format == PMX_IMG_F32) return im._imageData.f32.data();
        fbuf.push_back(as_float(im._imageData));
        return fbuf.back().data();
    };
    fbuf.reserve(allImage.size());
    const int esz = fmt == PMX_IMG_U8 ? 1 : 4;
    const Image& ref = allImage[refImageId];
    _refWidth = ref._imageData.width;
    _refHeight = ref._imageData.height;
    std::vector<int> sw, sh, spitch;
    std::vector<const void*> simg;
    std::vector<pmx_camera> scam;
    for (int i = 0; i < (int)allImage.size(); i++) {
        if (i == refImageId) continue;
        const GrayImage& g = allImage[i]._imageData;
        sw.push_back(g.width);
        sh.push_back(g.height);
        spitch.push_back(g.width * esz);
        simg.push_back(pixels(allImage[i]));
        scam.push_back(allImage[i].camera());
        _maxWidth = std::max(_maxWidth, g.width);
        _maxHeight = std::max(_maxHeight, g.height);
    }
    const pmx_camera rcam = ref.camera();
    const void* rimg = pixels(ref);
    PMX_SAFE_CALL(pmx_engine_create(pmx_dropin_context(), &p, fmt, _refWidth, _refHeight, rimg,
                                    _refWidth * esz, _numOfTargetImages, sw.data(), sh.data(),
                                    simg.data(), spitch.data(), &rcam, scam.data(), &_engine));
    PMX_SAFE_CALL(pmx_engine_prepare(_engine));
    std::vector<double> H((size_t)_numOfTargetImages * 18);
    PMX_SAFE_CALL(pmx_engine_get_homographies(_engine, H.data()));
    _transformHH = new float[H.size()];
    for (size_t i = 0; i < H.size(); i++) _transformHH[i] = (float)H[i];
    computeCUDAConfig(_refWidth, _refHeight, blockDim_x, blockDim_y);
    bindOutputs();
    /* _refImage and _psngState (src/patchMatch.cu:164-179) as views */
    float *I = nullptr, *sI = nullptr, *sII = nullptr;
    pmx_rng_state* st = nullptr;
    int rp = 0, sp = 0;
    PMX_SAFE_CALL(pmx_engine_export_state(_engine, &I, &sI, &sII, &rp, &st, &sp));
    auto view = [&](float* d) {
        return new Array2D_wrapper<float>(d, rp, _refWidth, _refHeight, 1, _blockDim_x, _blockDim_y);
    };
    _refImage = new Array2d_refImg(view(I), view(sI), view(sII), _refWidth, _refHeight);
    _psngState = new Array2D_psng(st, sp, _refWidth, _refHeight, _blockDim_x, _blockDim_y);
}

PatchMatch::~PatchMatch() {
    delete _refImage;
    delete _psngState;
    delete _depthMap;
    delete _matchCost;
    delete _usedImgsID;
    delete[] _transformHH;
    if (_engine) pmx_engine_destroy(_engine);
}

void PatchMatch::computeCUDAConfig(int width, int height, int blockDim_x, int blockDim_y) {
    _blockSize = dim3(blockDim_x, blockDim_y, 1);
    _gridSize = dim3((width - 1) / blockDim_x + 1, (height - 1) / blockDim_y + 1, 1);
}

void PatchMatch::bindOutputs() {
    float *d = nullptr, *c = nullptr;
    uint8_t* ids = nullptr;
    int dp = 0, cp = 0, ip = 0;
    PMX_SAFE_CALL(pmx_engine_export(_engine, &d, &dp, &c, &cp, &ids, &ip));
    delete _depthMap;
    delete _matchCost;
    delete _usedImgsID;
    _depthMap = new Array2D_wrapper<float>(d, dp, _refWidth, _refHeight, 1, _blockDim_x, _blockDim_y);
    _matchCost = new Array2D_wrapper<float>(c, cp, _refWidth, _refHeight, _numOfTargetImages,
                                            _blockDim_x, _blockDim_y);
    _usedImgsID = new Array2D_wrapper<uchar>(ids, ip, _refWidth, _refHeight, _numOfSamples,
                                             _blockDim_x, _blockDim_y);
}

void PatchMatch::runPatchMatch() {
    /* src/patchMatch.cu:263-270 (the constructor already limited hw to 1..22) */
    PMX_SAFE_CALL(pmx_engine_run(_engine));
    bindOutputs();
}
