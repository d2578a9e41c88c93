/*
 * pmx_io.cpp -- file formats of the reference program (see pmx_io.h).
 */
#include "pmx_io.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace {

bool set_err(std::string* e, const std::string& m) {
    if (e) *e = m;
    return false;
}

std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

/* OpenCV's BGR2GRAY for 8-bit data: Y = (4899 R + 9617 G + 1868 B + 2^13) >> 14 */
inline unsigned char gray_of(unsigned r, unsigned g, unsigned b) {
    return (unsigned char)((r * 4899u + g * 9617u + b * 1868u + 8192u) >> 14);
}

uint32_t be32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

inline int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

/* PNG (ISO/IEC 15948): non-interlaced, bit depth 8 or 16, colour types
 * 0, 2, 3 (palette), 4, 6.  16-bit samples keep their high byte, as
 * cv::imread's default 8-bit conversion does. */
bool read_png(const std::vector<unsigned char>& f, GrayImage& out, std::string* err) {
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return set_err(err, "not a PNG file");
    size_t pos = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<unsigned char> idat, plte;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        const std::string type(reinterpret_cast<const char*>(&f[pos + 4]), 4);
        if (pos + 12 + (size_t)len > f.size()) return set_err(err, "truncated PNG chunk");
        const unsigned char* d = &f[pos + 8];
        if (type == "IHDR") {
            if (len < 13) return set_err(err, "bad IHDR");
            W = be32(d);
            H = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (type == "PLTE") {
            plte.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (W == 0 || H == 0 || W > (1u << 20) || H > (1u << 20)) return set_err(err, "bad PNG size");
    if (interlace) return set_err(err, "interlaced PNG is not supported");
    if (depth != 8 && depth != 16) return set_err(err, "PNG bit depth must be 8 or 16");
    int ch = 0;
    switch (ctype) {
    case 0: ch = 1; break;
    case 2: ch = 3; break;
    case 3: ch = 1; break;
    case 4: ch = 2; break;
    case 6: ch = 4; break;
    default: return set_err(err, "unknown PNG colour type");
    }
    if (ctype == 3 && depth != 8) return set_err(err, "palette PNG must be 8-bit");
    const size_t bpp = (size_t)ch * (depth / 8);
    const size_t stride = bpp * W;
    std::vector<unsigned char> raw((stride + 1) * H);
    uLongf rawlen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK ||
        rawlen != raw.size())
        return set_err(err, "PNG data does not inflate to the image size");
    std::vector<unsigned char> img(stride * H), prev(stride, 0);
    for (uint32_t y = 0; y < H; y++) {
        const unsigned char* src = &raw[y * (stride + 1)];
        const int ft = src[0];
        unsigned char* row = &img[y * stride];
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= bpp ? row[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = src[1 + i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: return set_err(err, "bad PNG filter type");
            }
            row[i] = (unsigned char)v;
        }
        std::memcpy(prev.data(), row, stride);
    }
    out.width = (int)W;
    out.height = (int)H;
    out.format = PMX_IMG_U8;
    out.u8.resize((size_t)W * H);
    out.f32.clear();
    const size_t step = depth / 8; /* high byte first (big-endian samples) */
    for (size_t p = 0; p < (size_t)W * H; p++) {
        const unsigned char* s = &img[p * bpp];
        unsigned r, g, b;
        if (ctype == 3) {
            const size_t k = (size_t)s[0] * 3;
            if (k + 2 >= plte.size()) return set_err(err, "palette index out of range");
            r = plte[k];
            g = plte[k + 1];
            b = plte[k + 2];
        } else if (ch <= 2) {
            out.u8[p] = s[0];
            continue;
        } else {
            r = s[0];
            g = s[step];
            b = s[2 * step];
        }
        out.u8[p] = gray_of(r, g, b);
    }
    return true;
}

/* PGM/PPM, binary (P5/P6) or ASCII (P2/P3), maxval <= 65535 */
bool read_pnm(const std::vector<unsigned char>& f, GrayImage& out, std::string* err) {
    if (f.size() < 3 || f[0] != 'P' || (f[1] != '2' && f[1] != '3' && f[1] != '5' && f[1] != '6'))
        return set_err(err, "not a PGM/PPM file");
    const bool ascii = f[1] == '2' || f[1] == '3';
    const int ch = (f[1] == '3' || f[1] == '6') ? 3 : 1;
    size_t pos = 2;
    auto next_int = [&](long& v) -> bool {
        while (pos < f.size()) {
            if (f[pos] == '#') {
                while (pos < f.size() && f[pos] != '\n') pos++;
            } else if (std::isspace(f[pos])) {
                pos++;
            } else {
                break;
            }
        }
        if (pos >= f.size() || !std::isdigit(f[pos])) return false;
        v = 0;
        while (pos < f.size() && std::isdigit(f[pos])) v = v * 10 + (f[pos++] - '0');
        return true;
    };
    long W, H, maxv;
    if (!next_int(W) || !next_int(H) || !next_int(maxv) || W <= 0 || H <= 0 || maxv <= 0 ||
        maxv > 65535 || W > (1 << 20) || H > (1 << 20))
        return set_err(err, "bad PNM header");
    pos++; /* one whitespace byte before binary data */
    const size_t n = (size_t)W * H * ch;
    std::vector<unsigned> v(n);
    if (ascii) {
        for (size_t i = 0; i < n; i++) {
            long t;
            if (!next_int(t)) return set_err(err, "truncated PNM data");
            v[i] = (unsigned)t;
        }
    } else {
        const size_t bps = maxv > 255 ? 2 : 1;
        if (pos + n * bps > f.size()) return set_err(err, "truncated PNM data");
        for (size_t i = 0; i < n; i++)
            v[i] = bps == 2 ? ((unsigned)f[pos + 2 * i] << 8 | f[pos + 2 * i + 1]) : f[pos + i];
    }
    out.width = (int)W;
    out.height = (int)H;
    out.format = PMX_IMG_U8;
    out.u8.resize((size_t)W * H);
    out.f32.clear();
    auto to8 = [&](unsigned s) -> unsigned {
        return maxv == 255 ? s : (unsigned)std::lround((double)s * 255.0 / (double)maxv);
    };
    for (size_t p = 0; p < (size_t)W * H; p++)
        out.u8[p] = ch == 1 ? (unsigned char)to8(v[p])
                            : gray_of(to8(v[3 * p]), to8(v[3 * p + 1]), to8(v[3 * p + 2]));
    return true;
}

bool load_image(const std::string& name, GrayImage& img, std::string* err) {
    if (!pmx_read_gray(name, img, err)) {
        std::cout << "cannot read image " << name << ": " << (err ? *err : std::string()) << std::endl;
        return false;
    }
    return true;
}

} // namespace

bool pmx_read_ini(const std::string& fileName, std::map<std::string, std::string>& out,
                  std::string* error) {
    std::ifstream in(fileName);
    if (!in.is_open()) return set_err(error, "cannot open " + fileName);
    std::string line, section;
    int ln = 0;
    while (std::getline(in, line)) {
        ln++;
        std::string t = trim(line);
        if (t.empty() || t[0] == ';' || t[0] == '#') continue;
        if (t[0] == '[') {
            const size_t e = t.find(']');
            if (e == std::string::npos) return set_err(error, "bad section at line " + std::to_string(ln));
            section = trim(t.substr(1, e - 1));
            continue;
        }
        const size_t eq = t.find('=');
        if (eq == std::string::npos) return set_err(error, "expected key=value at line " + std::to_string(ln));
        if (section == "params") out[trim(t.substr(0, eq))] = trim(t.substr(eq + 1));
    }
    return true;
}

bool pmx_read_gray(const std::string& fileName, GrayImage& out, std::string* error) {
    std::ifstream in(fileName, std::ios::binary);
    if (!in.is_open()) return set_err(error, "cannot open " + fileName);
    std::vector<unsigned char> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (f.size() >= 8 && f[0] == 137 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G')
        return read_png(f, out, error);
    return read_pnm(f, out, error);
}

bool readMiddleBurry(std::string fileName, std::vector<Image>& allImages) {
    std::ifstream in(fileName);
    if (!in.is_open()) return false;
    unsigned int numOfImages = 0;
    if (!(in >> numOfImages) || numOfImages == 0) return false;
    const std::string dir = fileName.find('/') == std::string::npos
                                ? std::string()
                                : fileName.substr(0, fileName.rfind('/') + 1);
    allImages.resize(numOfImages);
    for (unsigned int i = 0; i < numOfImages; i++) {
        std::string imageFileName;
        float K[9], R[9], T[3];
        if (!(in >> imageFileName)) return false;
        for (int j = 0; j < 9; j++) in >> K[j];
        for (int j = 0; j < 9; j++) in >> R[j];
        for (int j = 0; j < 3; j++) in >> T[j];
        if (!in) return false;
        GrayImage img;
        std::string err;
        const std::string path = (imageFileName.size() && imageFileName[0] != '/') ? dir + imageFileName
                                                                                    : imageFileName;
        if (!load_image(path, img, &err)) return false;
        allImages[i].updateCamParam(K, R, T, img);
    }
    return true;
}

namespace {
/* src/image.cpp:94-118: unit quaternion (w, x, y, z) to a row-major rotation */
void quaternion_to_rotation(const double* q, double* R) {
    const double qq = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    double w = 1, x = 0, y = 0, z = 0;
    if (qq > 0) {
        w = q[0] / qq;
        x = q[1] / qq;
        y = q[2] / qq;
        z = q[3] / qq;
    }
    R[0] = w * w + x * x - z * z - y * y;
    R[1] = 2 * x * y - 2 * z * w;
    R[2] = 2 * y * w + 2 * z * x;
    R[3] = 2 * x * y + 2 * w * z;
    R[4] = y * y + w * w - z * z - x * x;
    R[5] = 2 * z * y - 2 * x * w;
    R[6] = 2 * x * z - 2 * y * w;
    R[7] = 2 * y * z + 2 * w * x;
    R[8] = z * z + w * w - y * y - x * x;
}
} // namespace

bool readNVM(std::string fileName, std::vector<Image>& allImages,
             std::vector<std::pair<float, float>>& depthRange) {
    std::ifstream in(fileName);
    if (!in.is_open() || in.peek() != 'N') return false;
    std::string token;
    in >> token;
    if (token.find("NVM_V3") == std::string::npos) return false;
    int ncam = 0;
    in >> ncam;
    if (ncam <= 1) return false;
    const std::string dir = fileName.find('/') == std::string::npos
                                ? std::string()
                                : fileName.substr(0, fileName.rfind('/') + 1);
    allImages.resize(ncam);
    for (int i = 0; i < ncam; i++) {
        std::string name;
        double f, q[4], c[3], d[2], R[9];
        in >> name >> f >> q[0] >> q[1] >> q[2] >> q[3] >> c[0] >> c[1] >> c[2] >> d[0] >> d[1];
        if (!in) return false;
        quaternion_to_rotation(q, R);
        GrayImage img;
        std::string err;
        if (!load_image(name.size() && name[0] != '/' ? dir + name : name, img, &err)) return false;
        float K[9] = {(float)f, 0.0f, (float)img.width / 2.0f, 0.0f, (float)f, img.height / 2.0f,
                      0.0f, 0.0f, 1.0f};
        float Rf[9], Tf[3];
        for (int j = 0; j < 9; j++) Rf[j] = (float)R[j];
        for (int j = 0; j < 3; j++) /* T = -R C */
            Tf[j] = (float)(-(R[j * 3] * c[0] + R[j * 3 + 1] * c[1] + R[j * 3 + 2] * c[2]));
        allImages[i].updateCamParam(K, Rf, Tf, img);
    }
    depthRange.assign(ncam, std::make_pair(FLT_MAX, -FLT_MAX));
    int npoint = 0;
    in >> npoint;
    if (npoint <= 0) return false;
    for (int i = 0; i < npoint; i++) {
        float pt[3];
        int rgb[3], npj;
        in >> pt[0] >> pt[1] >> pt[2] >> rgb[0] >> rgb[1] >> rgb[2] >> npj;
        if (!in) return false;
        for (int j = 0; j < npj; j++) {
            int cidx, fidx;
            float imx, imy;
            in >> cidx >> fidx >> imx >> imy;
            if (cidx < 0 || cidx >= ncam) continue;
            const float* P = allImages[cidx]._proj; /* z of K[R|T] (x,y,z,1) */
            const float z = P[8] * pt[0] + P[9] * pt[1] + P[10] * pt[2] + P[11];
            if (z <= 0) continue;
            depthRange[cidx].first = std::min(depthRange[cidx].first, z);
            depthRange[cidx].second = std::max(depthRange[cidx].second, z);
        }
    }
    for (auto& r : depthRange) { /* src/image.cpp:226-238 */
        const float size = r.second - r.first;
        float ratio = 0.25f;
        float nearv = r.first - ratio * size;
        while (nearv <= 0) {
            ratio /= 1.1f;
            nearv = r.first - ratio * size;
        }
        r.first = nearv;
        r.second += 0.25f * size;
    }
    std::cout << ncam << " cameras; " << npoint << " 3D points" << std::endl;
    return true;
}

bool pmx_read_depth_file(const std::string& fileName, int& width, int& height, int& depth,
                         std::vector<float>& data) {
    std::ifstream in(fileName, std::ios::binary);
    if (!in.is_open()) return false;
    char amp;
    if (!(in >> width >> amp) || amp != '&' || !(in >> height >> amp) || amp != '&' ||
        !(in >> depth >> amp) || amp != '&')
        return false;
    if (width <= 0 || height <= 0 || depth <= 0) return false;
    data.resize((size_t)width * height * depth);
    in.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(data.size() * sizeof(float)));
    return (size_t)in.gcount() == data.size() * sizeof(float);
}

/* ---- C entry points of the readers (used by the file-format tests) ---- */
extern "C" {

/* Decodes an image to 8-bit gray.  out may be NULL (size query); returns 0
 * on success, 1 on a decode error, 2 if cap is too small. */
int pmx_io_read_gray(const char* path, int* width, int* height, unsigned char* out, long cap) {
    GrayImage g;
    std::string err;
    if (!pmx_read_gray(path, g, &err)) return 1;
    *width = g.width;
    *height = g.height;
    if (!out) return 0;
    if (cap < (long)g.u8.size()) return 2;
    std::memcpy(out, g.u8.data(), g.u8.size());
    return 0;
}

/* Reads a camera list the way the reference program does (NVM_V3 first,
 * then Middlebury).  Returns 1 for NVM, 2 for Middlebury, 0 on failure.
 * Arrays hold cap cameras: K, R (9 floats each), T, C (3), image sizes,
 * and for NVM the widened depth range (near, far). */
int pmx_io_read_cameras(const char* path, int* n, float* K, float* R, float* T, float* C,
                        int* widths, int* heights, float* near_far, int cap) {
    std::vector<Image> imgs;
    std::vector<std::pair<float, float>> ranges;
    int kind = 0;
    if (readNVM(path, imgs, ranges)) kind = 1;
    else if (readMiddleBurry(path, imgs)) kind = 2;
    else return 0;
    *n = (int)imgs.size();
    for (int i = 0; i < (int)imgs.size() && i < cap; i++) {
        std::memcpy(K + 9 * i, imgs[i]._K, sizeof(float) * 9);
        std::memcpy(R + 9 * i, imgs[i]._R, sizeof(float) * 9);
        std::memcpy(T + 3 * i, imgs[i]._T, sizeof(float) * 3);
        std::memcpy(C + 3 * i, imgs[i]._C, sizeof(float) * 3);
        widths[i] = imgs[i]._imageData.width;
        heights[i] = imgs[i]._imageData.height;
        if (near_far && kind == 1) {
            near_far[2 * i] = ranges[i].first;
            near_far[2 * i + 1] = ranges[i].second;
        }
    }
    return kind;
}

/* Parses a depth-map file; data may be NULL (size query).  Returns 0 on
 * success. */
int pmx_io_read_depth_file(const char* path, int* width, int* height, int* depth, float* data,
                           long cap) {
    std::vector<float> v;
    if (!pmx_read_depth_file(path, *width, *height, *depth, v)) return 1;
    if (!data) return 0;
    if (cap < (long)v.size()) return 2;
    std::memcpy(data, v.data(), v.size() * sizeof(float));
    return 0;
}

} // extern "C"
