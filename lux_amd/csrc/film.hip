// film.hip -- film output at the end of the render path: the pixel bytes of
// write_ppm (src/film/ppm.h:37-56: scale by 1/spp, Reinhard c / (1 + c), clamp
// to [0, 1], 255.999 * c^(1/2.2) truncated) and the float body of write_pfm
// (src/film/pfm.h:27-44: scale by 1/spp, rows bottom to top), from a film of
// per-pixel sums resident in HBM, and whole-file writers on top of them.
//
// The gamma step is the only one whose result the GPU's powf could change. The
// byte is floor(255.999 * powf(c, 1/2.2)), a non-decreasing function of c, so
// it is the number of thresholds t_1..t_255 at or below c, where t_k is the
// smallest float in [0, 1] whose byte is at least k. The thresholds are found
// once on the host with the host's powf (the one the reference's writer calls),
// and the kernel counts them with an 8-step binary search. Every other step is
// a single IEEE operation, identical on both sides.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>

#include "lux_runtime.hpp"

namespace lux {

struct PpmTable {
    float t[256];  // t[1..255]; t[0] unused
};

namespace {

constexpr int kBlock = 256;

unsigned char ppm_byte_host(float c) { return static_cast<unsigned char>(255.999f * powf(c, 1.0f / 2.2f)); }

// Smallest float in [0, 1] whose byte reaches k (bisection over the float bit
// patterns, which are ordered like the values for non-negative floats).
float threshold(int k) {
    uint32_t lo = 0, hi = 0x3f800000u;  // byte(0) = 0 < k <= 255 = byte(1)
    while (lo + 1 < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        float c;
        std::memcpy(&c, &mid, sizeof(c));
        if (ppm_byte_host(c) >= k)
            hi = mid;
        else
            lo = mid;
    }
    float c;
    std::memcpy(&c, &hi, sizeof(c));
    return c;
}

const PpmTable& ppm_table() {
    static PpmTable table;
    static std::once_flag once;
    std::call_once(once, [] {
        table.t[0] = 0.0f;
        for (int k = 1; k < 256; ++k) table.t[k] = threshold(k);
    });
    return table;
}

// clamp01 of the reference (std::max(0, std::min(1, v))), NaN included:
// std::min(1, NaN) is 1 and std::max(0, x) keeps x unless 0 < x is false.
__device__ __forceinline__ float clamp01_ref(float v) {
    const float m = (v < 1.0f) ? v : 1.0f;
    return (0.0f < m) ? m : 0.0f;
}

__device__ __forceinline__ uint8_t ppm_byte(float pixel, float inv, const PpmTable& tab) {
    float c = pixel * inv;
    c = c / (1.0f + c);
    c = clamp01_ref(c);
    int lo = 0;  // count of thresholds t[1..255] <= c
#pragma unroll
    for (int step = 128; step > 0; step >>= 1)
        if (lo + step <= 255 && tab.t[lo + step] <= c) lo += step;
    return uint8_t(lo);
}

__global__ __launch_bounds__(kBlock) void k_film_ppm(const float* __restrict__ film, int64_t pixels, float inv,
                                                     PpmTable tab, uint8_t* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < 3 * pixels; i += int64_t(gridDim.x) * kBlock)
        out[i] = ppm_byte(film[i], inv, tab);
}

__global__ __launch_bounds__(kBlock) void k_film_pfm(const float* __restrict__ film, int width, int height,
                                                     float inv, float* __restrict__ out) {
    const int64_t n = int64_t(width) * height;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock) {
        const int64_t y = i / width, x = i - y * width;
        const int64_t src = (int64_t(height) - 1 - y) * width + x;  // output row y is film row h-1-y
        out[3 * i] = film[3 * src] * inv;
        out[3 * i + 1] = film[3 * src + 1] * inv;
        out[3 * i + 2] = film[3 * src + 2] * inv;
    }
}

float inv_sample_count(int spp) { return spp > 0 ? 1.0f / float(spp) : 0.0f; }

int grid_for(int64_t items) { return int(std::max<int64_t>(1, std::min<int64_t>(8192, (items + kBlock - 1) / kBlock))); }

void create_parent_directories(const std::string& path) {
    std::filesystem::path parent = std::filesystem::path(path).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent);
}

}  // namespace

void ppm_thresholds(float* out255) {
    const PpmTable& t = ppm_table();
    for (int k = 1; k < 256; ++k) out255[k - 1] = t.t[k];
}

void film_ppm(const float* film, int width, int height, int spp, uint8_t* out, hipStream_t s) {
    const int64_t pixels = int64_t(width) * height;
    if (pixels <= 0) return;
    LUX_LAUNCH("film_ppm", k_film_ppm, grid_for(3 * pixels), kBlock, 0, s, film, pixels, inv_sample_count(spp),
               ppm_table(), out);
}

void film_pfm(const float* film, int width, int height, int spp, float* out, hipStream_t s) {
    const int64_t pixels = int64_t(width) * height;
    if (pixels <= 0) return;
    LUX_LAUNCH("film_pfm", k_film_pfm, grid_for(pixels), kBlock, 0, s, film, width, height, inv_sample_count(spp),
               out);
}

void write_film_file(const char* path, bool ppm, const float* d_film, int width, int height, int spp, hipStream_t s) {
    const size_t pixels = size_t(width) * size_t(height);
    std::vector<char> body(ppm ? 3 * pixels : 12 * pixels);
    void* d_out = nullptr;
    if (!body.empty()) {
        hip_check(hipMalloc(&d_out, body.size()), "hipMalloc film output");
        try {
            if (ppm)
                film_ppm(d_film, width, height, spp, static_cast<uint8_t*>(d_out), s);
            else
                film_pfm(d_film, width, height, spp, static_cast<float*>(d_out), s);
            hip_check(hipMemcpyAsync(body.data(), d_out, body.size(), hipMemcpyDeviceToHost, s), "D2H film output");
            hip_check(hipStreamSynchronize(s), "synchronize");
        } catch (...) {
            (void)hipFree(d_out);
            throw;
        }
        hip_check(hipFree(d_out), "hipFree film output");
    }
    const std::string p(path);
    create_parent_directories(p);
    std::ofstream out(p, std::ios::binary);
    if (!out) throw Error(LUX_ERR_INVALID_ARGUMENT, "cannot open " + p);
    if (ppm)
        out << "P6\n" << width << " " << height << "\n255\n";
    else
        out << "PF\n" << width << " " << height << "\n-1.0\n";
    out.write(body.data(), std::streamsize(body.size()));
    if (!out) throw Error(LUX_ERR_HIP, "write failed: " + p);
}

}  // namespace lux
