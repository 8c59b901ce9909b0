// Drop-in replacement of include/cudaCompress/global.h (m0bl0/cudaCompress): basic typedefs.
#ifndef CUDACOMPRESS_AMD_DROPIN_GLOBAL_H
#define CUDACOMPRESS_AMD_DROPIN_GLOBAL_H

namespace cudaCompress {

typedef unsigned char byte;
typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
typedef long long int int64;
typedef unsigned long long int uint64;

}  // namespace cudaCompress

// The library exports a C ABI; the C++ API below is header-only, so no import/export decoration.
#ifndef CUCOMP_DLL
#define CUCOMP_DLL
#endif

#endif
