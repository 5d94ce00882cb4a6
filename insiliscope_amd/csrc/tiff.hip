// tiff.hip -- the movie file: little-endian baseline TIFF, one uncompressed grey page per frame (16-bit ADU or
// 32-bit float), classic or BigTIFF, as the reference CLI writes it (cli/tiff_writer.h). Host code; pages may come
// from device memory (a stack rendered on the GPU goes to the file without a caller-side copy).
//
// The file layout is the reference's, byte for byte (a format both sides must read): header; per page the sample
// data, then the NUL-terminated description (when given), a pad byte to an even offset, the IFD (classic: 2-byte
// count, 12-byte entries, 4-byte link; BigTIFF: 8-byte count, 20-byte entries, 8-byte link) with the tags
// 256 257 258 259 262 [270] 273 277 278 279 [339] in that order, and the previous IFD's link patched to it. A
// description that fits the entry's value field is stored in it. A file expected at or past 4 GiB less 64 MiB is
// a BigTIFF; a classic file that would pass 4 GiB is refused. A page may be written with its first rows only (the
// rest a hole of zeros) and completed row by row later.
#include "iscg_internal.h"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct iscg_tiff
{
   std::FILE* f = nullptr;
   std::string path, err;
   bool big = false;
   uint64_t link = 4;  // where the previous IFD's link (or the header's first-IFD offset) sits
   uint64_t end = 0;   // bytes in the file
   struct PageInfo
   {
      uint64_t data;
      uint32_t w, h;
   };
   std::vector<PageInfo> pages;
};

namespace {

constexpr uint64_t kClassicEnd = 0xFFFFFFFFull;

void fail_io(iscg_tiff* t, const char* what)
{
   const int e = errno;
   t->err = "TIFF " + t->path + ": " + what + " failed at page index " + std::to_string(t->pages.size()) +
            " after " + std::to_string(t->end) + " bytes (" + (e ? std::strerror(e) : "I/O error") + ")";
}

bool put(iscg_tiff* t, const void* p, size_t n)
{
   if (!t->err.empty())
      return false;
   if (n && std::fwrite(p, 1, n, t->f) != n)
   {
      fail_io(t, "write");
      return false;
   }
   t->end += n;
   return true;
}

bool seek(iscg_tiff* t, uint64_t at)
{
   if (fseeko(t->f, (off_t)at, SEEK_SET) != 0)
   {
      fail_io(t, "seek");
      return false;
   }
   return true;
}

// little-endian bytes of v, n of them, appended to b
void le(std::vector<uint8_t>& b, uint64_t v, int n)
{
   for (int i = 0; i < n; ++i)
      b.push_back((uint8_t)(v >> (8 * i)));
}

// One page: `given` bytes of px (the first rows; the rest a hole), the description, the IFD, the link to it.
int32_t write_page(iscg_tiff* t, const void* px, uint32_t bps, uint16_t sampleFormat, uint32_t w, uint32_t h,
                   const char* desc, uint64_t given)
{
   if (!t || !t->f || !t->err.empty())
      return -1;
   const uint64_t pageBytes = (uint64_t)w * h * bps;
   if (given > pageBytes)
      given = pageBytes;
   const size_t dlen = desc && desc[0] ? std::strlen(desc) : 0;
   const uint32_t descLen = dlen ? (uint32_t)dlen + 1 : 0;
   if (!t->big && t->end + pageBytes + descLen + 2 + 2 + 11 * 12 + 4 > kClassicEnd)
   {
      t->err = "TIFF " + t->path + ": page index " + std::to_string(t->pages.size()) +
               " ends beyond the 32-bit offsets of a classic TIFF; open the writer with bigTiff = 1 (or a larger "
               "expectedBytes) for files of this size";
      return -1;
   }
   errno = 0;
   const uint64_t dataOff = t->end;
   if (!put(t, px, (size_t)given))
      return -1;
   if (given < pageBytes)
   {
      t->end += pageBytes - given;
      if (!seek(t, t->end))
         return -1;
   }
   t->pages.push_back({dataOff, w, h});
   const uint64_t descOff = t->end;
   if (descLen && !put(t, desc, descLen))
      return -1;
   if (t->end & 1)
   {
      const uint8_t z = 0;
      if (!put(t, &z, 1))
         return -1;
   }
   const uint64_t ifd = t->end;
   // the entries: tag, type (2 ASCII, 3 SHORT, 4 LONG, 16 LONG8), count, value
   struct Entry
   {
      uint16_t tag, type;
      uint64_t count, value;
   };
   const uint16_t offType = t->big ? 16 : 4;
   std::vector<Entry> e = {{256, 4, 1, w}, {257, 4, 1, h}, {258, 3, 1, bps * 8}, {259, 3, 1, 1}, {262, 3, 1, 1}};
   if (descLen)
      e.push_back({270, 2, descLen, descOff});
   e.push_back({273, offType, 1, dataOff});
   e.push_back({277, 3, 1, 1});
   e.push_back({278, 4, 1, h});
   e.push_back({279, offType, 1, pageBytes});
   if (sampleFormat != 1)
      e.push_back({339, 3, 1, sampleFormat});
   const int valueBytes = t->big ? 8 : 4;
   std::vector<uint8_t> b;
   le(b, e.size(), t->big ? 8 : 2);
   for (const Entry& x : e)
   {
      le(b, x.tag, 2);
      le(b, x.type, 2);
      le(b, x.count, t->big ? 8 : 4);
      if (x.tag == 270 && descLen <= (uint32_t)valueBytes)
      {
         // a description that fits the value field is stored in it, zero-padded
         for (int i = 0; i < valueBytes; ++i)
            b.push_back(i < (int)descLen ? (uint8_t)desc[i] : 0);
      }
      else if (t->big && x.type == 3)
         le(b, x.value & 0xFFFF, 8);
      else if (t->big && x.type == 4)
         le(b, x.value & 0xFFFFFFFFull, 8);
      else
         le(b, x.value, valueBytes);
   }
   const uint64_t linkPos = t->end + b.size();
   le(b, 0, valueBytes);
   if (!put(t, b.data(), b.size()))
      return -1;
   // the previous IFD (or the header) points at this one
   std::vector<uint8_t> l;
   le(l, ifd, valueBytes);
   const uint64_t keep = t->end;
   if (!seek(t, t->link) || !put(t, l.data(), l.size()))
      return -1;
   t->end = keep;
   if (std::fseek(t->f, 0, SEEK_END) != 0)
   {
      fail_io(t, "seek");
      return -1;
   }
   if (std::fflush(t->f) != 0 || std::ferror(t->f))
   {
      fail_io(t, "write");
      return -1;
   }
   t->link = linkPos;
   return 0;
}

}  // namespace

extern "C" {

uint64_t iscg_tiff_bytes(uint64_t w, uint64_t h, uint64_t pages, uint64_t bytesPerSample, uint64_t descBytes)
{
   return 16 + pages * (w * h * bytesPerSample + 8 + 8 + 13 * 20 + 8) + descBytes + 2;
}

int32_t iscg_tiff_open(const char* path, uint64_t expectedBytes, int32_t bigTiff, iscg_tiff** out)
{
   if (!out)
      return -1;
   iscg_tiff* t = new iscg_tiff();
   *out = t;
   t->path = path ? path : "";
   t->big = bigTiff != 0 || expectedBytes + (64ull << 20) >= kClassicEnd;
   t->f = path ? std::fopen(path, "wb") : nullptr;
   if (!t->f)
   {
      t->err = "TIFF " + t->path + ": open for writing failed (" + std::strerror(errno) + ")";
      return -1;
   }
   const uint8_t classic[8] = {'I', 'I', 42, 0, 0, 0, 0, 0};
   const uint8_t bigHdr[16] = {'I', 'I', 43, 0, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
   t->link = t->big ? 8 : 4;
   return put(t, t->big ? bigHdr : classic, t->big ? 16 : 8) ? 0 : -1;
}

int32_t iscg_tiff_page_u16(iscg_tiff* t, const uint16_t* px, uint32_t w, uint32_t h, const char* desc)
{
   return write_page(t, px, 2, 1, w, h, desc, ~0ull);
}

int32_t iscg_tiff_page_f32(iscg_tiff* t, const float* px, uint32_t w, uint32_t h, const char* desc)
{
   return write_page(t, px, 4, 3, w, h, desc, ~0ull);
}

int32_t iscg_tiff_page_rows_u16(iscg_tiff* t, const uint16_t* px, uint32_t rows, uint32_t w, uint32_t h,
                                const char* desc)
{
   return write_page(t, px, 2, 1, w, h, desc, (uint64_t)(rows < h ? rows : h) * w * 2);
}

int32_t iscg_tiff_rows_u16(iscg_tiff* t, uint64_t page, uint32_t y0, uint32_t rows, const uint16_t* px)
{
   if (!t || !t->f || !t->err.empty())
      return -1;
   if (page >= t->pages.size() || (uint64_t)y0 + rows > t->pages[page].h)
   {
      t->err = "TIFF " + t->path + ": rows [" + std::to_string(y0) + ", " + std::to_string(y0 + rows) +
               ") lie outside page index " + std::to_string(page) + " (or the page does not exist)";
      return -1;
   }
   const uint64_t keep = t->end, row = (uint64_t)t->pages[page].w * 2;
   errno = 0;
   if (!seek(t, t->pages[page].data + y0 * row) || !put(t, px, (size_t)(rows * row)))
      return -1;
   t->end = keep;
   if (!seek(t, t->end))
      return -1;
   if (std::fflush(t->f) != 0 || std::ferror(t->f))
   {
      fail_io(t, "write");
      return -1;
   }
   return 0;
}

int32_t iscg_tiff_frames_u16(iscg_tiff* t, const uint16_t* frames, uint32_t nFrames, uint32_t w, uint32_t h,
                             const char* firstDesc)
{
   if (!t)
      return -1;
   const size_t n = (size_t)w * h;
   const bool dev = iscg::is_device_ptr(frames);
   std::vector<uint16_t> host(dev ? n : 0);
   for (uint32_t f = 0; f < nFrames; ++f)
   {
      const uint16_t* src = frames + (size_t)f * n;
      if (dev)
      {
         const hipError_t e = hipMemcpy(host.data(), src, n * sizeof(uint16_t), hipMemcpyDeviceToHost);
         if (e != hipSuccess)
         {
            t->err =
               std::string("TIFF: device-to-host copy of frame ") + std::to_string(f) + " failed (" +
               hipGetErrorString(e) + ")";
            return -1;
         }
         src = host.data();
      }
      if (write_page(t, src, 2, 1, w, h, f == 0 ? firstDesc : nullptr, ~0ull))
         return -1;
   }
   return 0;
}

int32_t iscg_tiff_info(const iscg_tiff* t, uint64_t* pages, uint64_t* bytes, int32_t* big)
{
   if (!t)
      return -1;
   if (pages)
      *pages = t->pages.size();
   if (bytes)
      *bytes = t->end;
   if (big)
      *big = t->big ? 1 : 0;
   return 0;
}

const char* iscg_tiff_error(const iscg_tiff* t) { return t ? t->err.c_str() : "null TIFF writer"; }

int32_t iscg_tiff_close(iscg_tiff* t)
{
   if (!t)
      return -1;
   if (t->f)
   {
      errno = 0;
      const bool closed = std::fclose(t->f) == 0;
      t->f = nullptr;
      if (!closed && t->err.empty())
         fail_io(t, "close");
   }
   return t->err.empty() ? 0 : -1;
}

void iscg_tiff_destroy(iscg_tiff* t)
{
   if (t && t->f)
      std::fclose(t->f);
   delete t;
}

}  // extern "C"
