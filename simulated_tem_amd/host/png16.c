/*
 * png16.c -- 16-bit greyscale PNG writer for the d2bi host (zlib only).
 *
 * Pixel encoding follows write_greyscale_png_file (src/Lib_Greyscale.c:162):
 * value f clamped to [0, 0.999984741], stored as (int)(f*65536) big-endian,
 * one row per image line, row 0 first.  The PNG container (signature, IHDR,
 * one IDAT of zlib-deflated filter-0 rows, IEND, CRC-32 per chunk) is
 * written here directly.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

static void put32(unsigned char* p, uint32_t v)
{
    p[0] = (unsigned char)(v >> 24); p[1] = (unsigned char)(v >> 16);
    p[2] = (unsigned char)(v >> 8);  p[3] = (unsigned char)v;
}

static int chunk(FILE* fp, const char* type, const unsigned char* data, uint32_t len)
{
    unsigned char hdr[8];
    put32(hdr, len);
    memcpy(hdr + 4, type, 4);
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, (const Bytef*)type, 4);
    if (len) crc = crc32(crc, data, len);
    unsigned char tail[4];
    put32(tail, (uint32_t)crc);
    if (fwrite(hdr, 1, 8, fp) != 8) return -1;
    if (len && fwrite(data, 1, len, fp) != len) return -1;
    if (fwrite(tail, 1, 4, fp) != 4) return -1;
    return 0;
}

/* img: width*height floats, x fastest.  Returns 0 on success. */
int d2bi_write_png16(const char* path, int width, int height, const float* img)
{
    if (!path || !img || width < 1 || height < 1) return -1;
    const size_t row = 1 + 2 * (size_t)width;
    const size_t raw_len = row * (size_t)height;
    unsigned char* raw = (unsigned char*)malloc(raw_len);
    if (!raw) return -2;
    for (int y = 0; y < height; y++) {
        unsigned char* r = raw + row * y;
        r[0] = 0;                                   /* filter type 0 */
        for (int x = 0; x < width; x++) {
            float f = img[(size_t)x + (size_t)width * y];
            if (f > 0.999984741f) f = 0.999984741f;
            else if (f < 0.0f) f = 0.0f;
            int i = (int)(f * 65536);
            r[1 + 2 * x] = (unsigned char)(i >> 8);
            r[2 + 2 * x] = (unsigned char)(i % 256);
        }
    }
    uLongf zlen = compressBound((uLong)raw_len);
    unsigned char* z = (unsigned char*)malloc(zlen);
    if (!z) { free(raw); return -2; }
    if (compress2(z, &zlen, raw, (uLong)raw_len, 6) != Z_OK) { free(raw); free(z); return -3; }
    free(raw);
    FILE* fp = fopen(path, "wb");
    if (!fp) { free(z); return -4; }
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    unsigned char ihdr[13];
    put32(ihdr, (uint32_t)width);
    put32(ihdr + 4, (uint32_t)height);
    ihdr[8] = 16;    /* bit depth */
    ihdr[9] = 0;     /* greyscale */
    ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0;
    int rc = 0;
    if (fwrite(sig, 1, 8, fp) != 8) rc = -5;
    if (!rc) rc = chunk(fp, "IHDR", ihdr, 13);
    if (!rc) rc = chunk(fp, "IDAT", z, (uint32_t)zlen);
    if (!rc) rc = chunk(fp, "IEND", NULL, 0);
    fclose(fp);
    free(z);
    return rc;
}
