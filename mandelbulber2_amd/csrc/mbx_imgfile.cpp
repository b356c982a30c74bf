// mbx_imgfile.cpp - TIFF and OpenEXR output of the image and its channels, the other two
// writers of the reference's image saving (mandelbulber2/src/file_image.cpp): the files are
// assembled here from the published formats, with zlib for the deflate streams.
//
// TIFF (ImageFileSaveTIFF::SaveTIFF, file_image.cpp:1590-1760): baseline little-endian TIFF with
// the fields the reference sets through libtiff: 8/16-bit unsigned or 32-bit IEEE float samples,
// 1 (grey, min-is-black), 3 (RGB) or 4 (RGB + alpha) samples per pixel, contiguous planar
// configuration, MSB-to-LSB fill order, strips of 64 rows (SAVE_CHUNK_SIZE, file_image.hpp:150)
// deflated with the legacy deflate compression code 32946 (COMPRESSION_DEFLATE).
//
// OpenEXR (ImageFileSaveEXR::SaveEXR, file_image.cpp:1249-1530): a single-part scanline file
// with ZIPS compression (one scanline per chunk, Imf::ZIPS_COMPRESSION), HALF or FLOAT channels;
// the header carries the channel list (sorted by name, as the format requires), compression,
// data and display windows, line order, pixel aspect ratio and screen window.  A compressed chunk
// is the zlib stream of the scanline's bytes after the format's byte interleave and delta
// predictor; a chunk that would not shrink is stored raw, as the format allows.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <zlib.h>

#include "../../include/mbx.h"

namespace
{

void le16(std::vector<uint8_t> &b, uint16_t v)
{
	b.push_back((uint8_t)v);
	b.push_back((uint8_t)(v >> 8));
}
void le32(std::vector<uint8_t> &b, uint32_t v)
{
	for (int k = 0; k < 4; k++) b.push_back((uint8_t)(v >> (8 * k)));
}
void le64(std::vector<uint8_t> &b, uint64_t v)
{
	for (int k = 0; k < 8; k++) b.push_back((uint8_t)(v >> (8 * k)));
}
void set32(std::vector<uint8_t> &b, size_t at, uint32_t v)
{
	for (int k = 0; k < 4; k++) b[at + k] = (uint8_t)(v >> (8 * k));
}
void set64(std::vector<uint8_t> &b, size_t at, uint64_t v)
{
	for (int k = 0; k < 8; k++) b[at + k] = (uint8_t)(v >> (8 * k));
}

bool deflate_all(const uint8_t *src, size_t n, std::vector<uint8_t> &out)
{
	uLongf len = compressBound((uLong)n);
	out.resize(len);
	if (compress2(out.data(), &len, src, (uLong)n, Z_DEFAULT_COMPRESSION) != Z_OK) return false;
	out.resize(len);
	return true;
}

bool write_file(const char *path, const std::vector<uint8_t> &file)
{
	FILE *fp = std::fopen(path, "wb");
	if (!fp) return false;
	const bool ok = std::fwrite(file.data(), 1, file.size(), fp) == file.size();
	return (std::fclose(fp) == 0) && ok;
}

// IEEE binary16 of a float, rounded to nearest with ties to even; overflow to infinity, NaN kept
// a NaN (the conversion OpenEXR's half type performs)
uint16_t float_to_half(float f)
{
	uint32_t x;
	std::memcpy(&x, &f, 4);
	const uint32_t sign = (x >> 16) & 0x8000u;
	const uint32_t absx = x & 0x7fffffffu;
	if (absx >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u : 0u));
	if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); // rounds to >= 65520: infinity
	if (absx < 0x38800000u)
	{
		// subnormal half (or zero): value = m * 2^-24 with m the rounded integer
		if (absx < 0x33000000u) return (uint16_t)sign; // below half the smallest subnormal
		const uint32_t e = absx >> 23;
		const uint32_t mant = (absx & 0x7fffffu) | 0x800000u;
		const uint32_t shift = 126u - e; // 14 .. 24
		uint32_t m = mant >> shift;
		const uint32_t rem = mant & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
		if (rem > halfway || (rem == halfway && (m & 1u))) m++;
		return (uint16_t)(sign | m);
	}
	uint32_t h = ((absx >> 13) - (112u << 10)); // rebias the exponent 127 -> 15
	const uint32_t rem = absx & 0x1fffu;
	if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
	return (uint16_t)(sign | h);
}

} // namespace

extern "C" uint16_t mbx_float_to_half(float f) { return float_to_half(f); }

extern "C" int mbx_write_tiff(const char *path, int32_t width, int32_t height, int32_t bits, int32_t channels,
	const void *pixels)
{
	if (!path || !pixels || width <= 0 || height <= 0 || (bits != 8 && bits != 16 && bits != 32) ||
		(channels != 1 && channels != 3 && channels != 4))
		return MBX_ERR_INVALID;
	const size_t bps = (size_t)bits / 8;
	const size_t rowBytes = (size_t)width * (size_t)channels * bps;
	if (rowBytes / bps / (size_t)channels != (size_t)width || (uint64_t)rowBytes * (uint64_t)height > 0xffffffffull)
		return MBX_ERR_INVALID;
	const uint32_t rowsPerStrip = 64;
	const uint32_t strips = (uint32_t)((height + rowsPerStrip - 1) / rowsPerStrip);
	std::vector<uint8_t> file;
	file.push_back('I');
	file.push_back('I');
	le16(file, 42);
	le32(file, 0); // offset of the IFD, set below
	std::vector<uint32_t> offsets(strips), counts(strips);
	const uint8_t *src = (const uint8_t *)pixels;
	for (uint32_t s = 0; s < strips; s++)
	{
		const uint32_t y0 = s * rowsPerStrip;
		const uint32_t rows = std::min<uint32_t>(rowsPerStrip, (uint32_t)height - y0);
		std::vector<uint8_t> z;
		if (!deflate_all(src + (size_t)y0 * rowBytes, (size_t)rows * rowBytes, z)) return MBX_ERR_INVALID;
		offsets[s] = (uint32_t)file.size();
		counts[s] = (uint32_t)z.size();
		file.insert(file.end(), z.begin(), z.end());
		if (file.size() & 1u) file.push_back(0); // word alignment of what follows
	}
	// out-of-line field values: bits per sample, sample format, strip offsets and byte counts
	auto array16 = [&](uint16_t v, int n) {
		const uint32_t at = (uint32_t)file.size();
		for (int k = 0; k < n; k++) le16(file, v);
		if (file.size() & 1u) file.push_back(0);
		return at;
	};
	const uint16_t sampleFormat = bits == 32 ? 3 : 1; // IEEE float / unsigned integer
	const uint32_t bitsAt = channels > 2 ? array16((uint16_t)bits, channels) : 0;
	const uint32_t formatAt = channels > 2 ? array16(sampleFormat, channels) : 0;
	uint32_t offsetsAt = 0, countsAt = 0;
	if (strips > 1)
	{
		offsetsAt = (uint32_t)file.size();
		for (uint32_t v : offsets) le32(file, v);
		countsAt = (uint32_t)file.size();
		for (uint32_t v : counts) le32(file, v);
	}
	struct Field
	{
		uint16_t tag, type;
		uint32_t count, value;
	};
	// types: 3 SHORT, 4 LONG; a SHORT value sits in the low half of the value field
	std::vector<Field> f;
	f.push_back({256, 4, 1, (uint32_t)width});
	f.push_back({257, 4, 1, (uint32_t)height});
	f.push_back({258, 3, (uint32_t)channels, channels > 2 ? bitsAt : (uint32_t)bits});
	f.push_back({259, 3, 1, 32946u});                   // deflate (COMPRESSION_DEFLATE)
	f.push_back({262, 3, 1, channels >= 3 ? 2u : 1u});  // RGB / min-is-black
	f.push_back({266, 3, 1, 1u});                       // fill order MSB to LSB
	f.push_back({273, 4, strips, strips > 1 ? offsetsAt : offsets[0]});
	f.push_back({277, 3, 1, (uint32_t)channels});
	f.push_back({278, 4, 1, rowsPerStrip});
	f.push_back({279, 4, strips, strips > 1 ? countsAt : counts[0]});
	f.push_back({284, 3, 1, 1u});                       // contiguous samples
	if (channels == 4) f.push_back({338, 3, 1, 2u});    // the fourth sample: unassociated alpha
	f.push_back({339, 3, (uint32_t)channels, channels > 2 ? formatAt : (uint32_t)sampleFormat});
	const uint32_t ifd = (uint32_t)file.size();
	set32(file, 4, ifd);
	le16(file, (uint16_t)f.size());
	for (const Field &e : f)
	{
		le16(file, e.tag);
		le16(file, e.type);
		le32(file, e.count);
		if (e.type == 3 && e.count == 1)
		{
			le16(file, (uint16_t)e.value);
			le16(file, 0);
		}
		else if (e.type == 3 && e.count == 2)
			return MBX_ERR_INVALID; // (not produced: 2 samples per pixel are refused above)
		else
			le32(file, e.value);
	}
	le32(file, 0); // no further IFD
	return write_file(path, file) ? MBX_OK : MBX_ERR_IO;
}

extern "C" int mbx_write_exr(const char *path, int32_t width, int32_t height, int32_t nchannels,
	const char *const *names, const int32_t *pixel_types, const float *const *planes, const int32_t *strides,
	int32_t linear)
{
	if (!path || !names || !pixel_types || !planes || !strides || width <= 0 || height <= 0 || nchannels < 1 ||
		nchannels > 64)
		return MBX_ERR_INVALID;
	struct Ch
	{
		std::string name;
		int type; // 1 HALF, 2 FLOAT
		const float *plane;
		int stride;
	};
	std::vector<Ch> ch;
	for (int c = 0; c < nchannels; c++)
	{
		if (!names[c] || !planes[c] || (pixel_types[c] != 1 && pixel_types[c] != 2) || strides[c] < 1) return MBX_ERR_INVALID;
		const size_t len = std::strlen(names[c]);
		if (len < 1 || len > 31) return MBX_ERR_INVALID; // short names (no long-name flag)
		ch.push_back({names[c], pixel_types[c], planes[c], strides[c]});
	}
	std::sort(ch.begin(), ch.end(), [](const Ch &a, const Ch &b) { return a.name < b.name; });
	for (size_t c = 1; c < ch.size(); c++)
		if (ch[c].name == ch[c - 1].name) return MBX_ERR_INVALID;
	std::vector<uint8_t> file;
	le32(file, 20000630u); // magic
	le32(file, 2u);        // version 2, single-part scanline, short names
	auto attr = [&](const char *name, const char *type, const std::vector<uint8_t> &value) {
		file.insert(file.end(), name, name + std::strlen(name) + 1);
		file.insert(file.end(), type, type + std::strlen(type) + 1);
		le32(file, (uint32_t)value.size());
		file.insert(file.end(), value.begin(), value.end());
	};
	std::vector<uint8_t> v;
	for (const Ch &c : ch)
	{
		v.insert(v.end(), c.name.begin(), c.name.end());
		v.push_back(0);
		le32(v, (uint32_t)c.type);
		v.push_back(linear ? 1 : 0); // pLinear
		v.push_back(0);
		v.push_back(0);
		v.push_back(0);
		le32(v, 1); // x sampling
		le32(v, 1); // y sampling
	}
	v.push_back(0);
	attr("channels", "chlist", v);
	attr("compression", "compression", std::vector<uint8_t>{2}); // ZIPS
	v.clear();
	le32(v, 0);
	le32(v, 0);
	le32(v, (uint32_t)(width - 1));
	le32(v, (uint32_t)(height - 1));
	attr("dataWindow", "box2i", v);
	attr("displayWindow", "box2i", v);
	attr("lineOrder", "lineOrder", std::vector<uint8_t>{0}); // increasing y
	const float one = 1.0f, zero = 0.0f;
	v.assign((const uint8_t *)&one, (const uint8_t *)&one + 4);
	attr("pixelAspectRatio", "float", v);
	v.assign((const uint8_t *)&zero, (const uint8_t *)&zero + 4);
	v.insert(v.end(), (const uint8_t *)&zero, (const uint8_t *)&zero + 4);
	attr("screenWindowCenter", "v2f", v);
	v.assign((const uint8_t *)&one, (const uint8_t *)&one + 4);
	attr("screenWindowWidth", "float", v);
	file.push_back(0); // end of header
	const size_t tableAt = file.size();
	for (int y = 0; y < height; y++) le64(file, 0);
	size_t lineBytes = 0;
	for (const Ch &c : ch) lineBytes += (size_t)width * (c.type == 1 ? 2u : 4u);
	std::vector<uint8_t> raw(lineBytes), tmp(lineBytes), z;
	for (int y = 0; y < height; y++)
	{
		size_t at = 0;
		for (const Ch &c : ch)
			for (int x = 0; x < width; x++)
			{
				const float val = c.plane[((size_t)y * (size_t)width + (size_t)x) * (size_t)c.stride];
				if (c.type == 1)
				{
					const uint16_t h = float_to_half(val);
					raw[at++] = (uint8_t)h;
					raw[at++] = (uint8_t)(h >> 8);
				}
				else
				{
					std::memcpy(&raw[at], &val, 4);
					at += 4;
				}
			}
		// the format's zip filter: bytes split into even- and odd-indexed halves, then each byte
		// replaced by its difference from the previous one, offset by 128
		size_t i1 = 0, i2 = (lineBytes + 1) / 2;
		for (size_t k = 0; k < lineBytes; k++) tmp[(k & 1u) ? i2++ : i1++] = raw[k];
		for (size_t k = lineBytes - 1; k > 0; k--) tmp[k] = (uint8_t)((int)tmp[k] - (int)tmp[k - 1] + 128);
		if (!deflate_all(tmp.data(), lineBytes, z)) return MBX_ERR_INVALID;
		set64(file, tableAt + 8 * (size_t)y, file.size());
		le32(file, (uint32_t)y);
		if (z.size() < lineBytes)
		{
			le32(file, (uint32_t)z.size());
			file.insert(file.end(), z.begin(), z.end());
		}
		else
		{
			le32(file, (uint32_t)lineBytes);
			file.insert(file.end(), raw.begin(), raw.end());
		}
	}
	return write_file(path, file) ? MBX_OK : MBX_ERR_IO;
}
