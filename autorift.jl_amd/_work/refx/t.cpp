#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <random>
static uint32_t alignbyte(uint32_t hi, uint32_t lo, uint32_t s) { return (uint32_t)(((((uint64_t)hi) << 32) | lo) >> (8 * (s & 3))); }
int main() {
    std::mt19937 rng(1);
    int bad = 0;
    for (int64_t n : {1, 3, 15, 16, 17, 31, 32, 33, 100, 1000, 4096, 4097, 10007}) {
        int64_t stride = (n + 1024 + 255) / 256 * 256;
        std::vector<uint8_t> src(n + 64), ref(4 * stride), got(4 * stride);
        for (int64_t i = 0; i < n; i++) src[i] = rng();
        for (int64_t y = 0; y < stride; y++) for (int k = 0; k < 4; k++) ref[k * stride + y] = (y >= k && y - k < n) ? (uint8_t)(src[y - k] ^ 0x80) : 0x80;
        for (int64_t v = 0; v < stride / 16; v++) {
            int64_t y0 = v * 16; uint32_t o[4][4];
            if (y0 >= 16 && y0 + 16 <= n) {
                uint32_t c[5]; memcpy(&c[0], &src[y0 - 4], 4); memcpy(&c[1], &src[y0], 16);
                for (int d = 0; d < 4; d++) { o[0][d] = c[d + 1] ^ 0x80808080u; for (int k = 1; k < 4; k++) o[k][d] = alignbyte(c[d + 1], c[d], 4 - k) ^ 0x80808080u; }
            } else {
                for (int k = 0; k < 4; k++) for (int d = 0; d < 4; d++) { uint32_t w = 0; for (int i = 0; i < 4; i++) { int64_t y = y0 + 4 * d + i; uint32_t b = (y >= k && y - k < n) ? (uint32_t)(src[y - k] ^ 0x80) : 0x80u; w |= b << (8 * i);} o[k][d] = w; }
            }
            for (int k = 0; k < 4; k++) memcpy(&got[k * stride + y0], o[k], 16);
        }
        if (memcmp(ref.data(), got.data(), ref.size())) { printf("mismatch n=%ld\n", (long)n); bad++; }
    }
    printf(bad ? "FAIL\n" : "all equal\n");
    return bad;
}
