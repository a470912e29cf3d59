// film_oracle.cpp -- CPU restatement of the reference's film writers
// (write_ppm, src/film/ppm.h:37-56; write_pfm, src/film/pfm.h:27-44) and an
// exhaustive check of a gamma threshold table against the byte formula.
// TEST INFRASTRUCTURE ONLY.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {
float clamp01(float v) {  // std::max(0, std::min(1, v))
    const float m = (v < 1.0f) ? v : 1.0f;
    return (0.0f < m) ? m : 0.0f;
}
unsigned char ppm_byte(float pixel, float inv) {
    float c = pixel * inv;
    c = c / (1.0f + c);
    c = clamp01(c);
    return static_cast<unsigned char>(255.999f * powf(c, 1.0f / 2.2f));
}
float inv_count(int spp) { return spp > 0 ? 1.0f / float(spp) : 0.0f; }
}  // namespace

extern "C" {

// The whole file: "P6\n<w> <h>\n255\n" + 3 bytes per pixel in film order.
int64_t orc_ppm_file(const float* film, int32_t w, int32_t h, int32_t spp, uint8_t* out, int64_t cap) {
    const std::string head = "P6\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
    const int64_t n = int64_t(head.size()) + 3 * int64_t(w) * h;
    if (out && cap >= n) {
        std::memcpy(out, head.data(), head.size());
        const float inv = inv_count(spp);
        for (int64_t i = 0; i < 3 * int64_t(w) * h; ++i) out[head.size() + i] = ppm_byte(film[i], inv);
    }
    return n;
}

// The whole file: "PF\n<w> <h>\n-1.0\n" + rows from y = h-1 down to 0.
int64_t orc_pfm_file(const float* film, int32_t w, int32_t h, int32_t spp, uint8_t* out, int64_t cap) {
    const std::string head = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
    const int64_t n = int64_t(head.size()) + 12 * int64_t(w) * h;
    if (out && cap >= n) {
        std::memcpy(out, head.data(), head.size());
        const float inv = inv_count(spp);
        uint8_t* p = out + head.size();
        for (int y = h - 1; y >= 0; --y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c) {
                    const float v = film[3 * (int64_t(y) * w + x) + c] * inv;
                    std::memcpy(p, &v, 4);
                    p += 4;
                }
    }
    return n;
}

// Floats c in [0, 1] where the count of thresholds t[0..254] <= c differs from
// the byte (unsigned char)(255.999f * powf(c, 1.0f / 2.2f)): every bit pattern
// from 0 to 0x3f800000, split over `threads` threads.
int64_t orc_ppm_table_mismatches(const float* t255, int32_t threads) {
    const uint32_t end = 0x3f800000u;
    std::vector<int64_t> bad(threads > 0 ? threads : 1, 0);
    std::vector<std::thread> pool;
    const int nt = int(bad.size());
    for (int k = 0; k < nt; ++k)
        pool.emplace_back([&, k] {
            for (uint32_t b = uint32_t(k); b <= end; b += uint32_t(nt)) {
                float c;
                std::memcpy(&c, &b, 4);
                const int ref = static_cast<unsigned char>(255.999f * powf(c, 1.0f / 2.2f));
                int lo = 0;
                for (int step = 128; step > 0; step >>= 1)
                    if (lo + step <= 255 && t255[lo + step - 1] <= c) lo += step;
                bad[k] += lo != ref;
            }
        });
    for (auto& t : pool) t.join();
    int64_t total = 0;
    for (int64_t v : bad) total += v;
    return total;
}

}  // extern "C"
