// mbx_png.cpp - PNG output of the compiled image (ImageFileSavePNG::SavePNG,
// mandelbulber2/src/file_image.cpp:481-600): 8- or 16-bit samples, RGB (colour image) or grey
// (z-buffer image), not interlaced, 16-bit samples big-endian as png_set_swap stores them.  The
// file is assembled here (signature, IHDR, one IDAT with the zlib stream of the scanlines, IEND)
// and compressed with zlib at its default level; every scanline uses filter type 0.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include <zlib.h>

#include "../../include/mbx.h"

namespace
{

void put32(std::vector<uint8_t> &b, uint32_t v)
{
	b.push_back((uint8_t)(v >> 24));
	b.push_back((uint8_t)(v >> 16));
	b.push_back((uint8_t)(v >> 8));
	b.push_back((uint8_t)v);
}

// length, type, data, CRC-32 of type + data
void chunk(std::vector<uint8_t> &file, const char *type, const uint8_t *data, size_t len)
{
	put32(file, (uint32_t)len);
	const size_t at = file.size();
	file.insert(file.end(), type, type + 4);
	if (len) file.insert(file.end(), data, data + len);
	const uLong crc = crc32(0L, file.data() + at, (uInt)(len + 4));
	put32(file, (uint32_t)crc);
}

} // namespace

extern "C" int mbx_write_png(const char *path, int32_t width, int32_t height, int32_t bits, int32_t channels,
	const void *pixels)
{
	if (!path || !pixels || width <= 0 || height <= 0 || (bits != 8 && bits != 16) ||
		(channels != 1 && channels != 3))
		return MBX_ERR_INVALID;
	const size_t bytesPerSample = (size_t)bits / 8;
	const size_t rowBytes = (size_t)width * (size_t)channels * bytesPerSample;
	if (rowBytes / bytesPerSample / (size_t)channels != (size_t)width) return MBX_ERR_INVALID;
	std::vector<uint8_t> raw;
	raw.resize((rowBytes + 1) * (size_t)height);
	const uint8_t *src8 = (const uint8_t *)pixels;
	const uint16_t *src16 = (const uint16_t *)pixels;
	const size_t samplesPerRow = (size_t)width * (size_t)channels;
	for (size_t y = 0; y < (size_t)height; y++)
	{
		uint8_t *row = raw.data() + y * (rowBytes + 1);
		row[0] = 0; // filter type None
		if (bits == 8)
			std::memcpy(row + 1, src8 + y * samplesPerRow, rowBytes);
		else
			for (size_t s = 0; s < samplesPerRow; s++)
			{
				const uint16_t v = src16[y * samplesPerRow + s];
				row[1 + 2 * s] = (uint8_t)(v >> 8);
				row[2 + 2 * s] = (uint8_t)v;
			}
	}
	uLongf zlen = compressBound((uLong)raw.size());
	std::vector<uint8_t> z(zlen);
	if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
		return MBX_ERR_NOMEM;
	std::vector<uint8_t> file = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
	std::vector<uint8_t> ihdr;
	put32(ihdr, (uint32_t)width);
	put32(ihdr, (uint32_t)height);
	ihdr.push_back((uint8_t)bits);
	ihdr.push_back(channels == 3 ? 2 : 0); // colour type RGB or grey
	ihdr.push_back(0);                     // deflate
	ihdr.push_back(0);                     // adaptive filtering (filter method 0)
	ihdr.push_back(0);                     // not interlaced
	chunk(file, "IHDR", ihdr.data(), ihdr.size());
	chunk(file, "IDAT", z.data(), zlen);
	chunk(file, "IEND", nullptr, 0);
	FILE *fp = std::fopen(path, "wb");
	if (!fp) return MBX_ERR_IO;
	const bool ok = std::fwrite(file.data(), 1, file.size(), fp) == file.size();
	return (std::fclose(fp) == 0 && ok) ? MBX_OK : MBX_ERR_IO;
}
