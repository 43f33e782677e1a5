// Readable element-type names used in benchmark names ("fp32", ...), as in the
// reference's include/type_traits.h:4-23.
#pragma once

template <class T>
struct ReadableTypename;

#define MFX_READABLE_TYPENAME(T, str)                                                                                  \
  template <>                                                                                                          \
  struct ReadableTypename<T> {                                                                                         \
    static constexpr const char* name = str;                                                                           \
  };

MFX_READABLE_TYPENAME(char, "int8")
MFX_READABLE_TYPENAME(unsigned char, "uint8")
MFX_READABLE_TYPENAME(short, "int16")
MFX_READABLE_TYPENAME(unsigned short, "uint16")
MFX_READABLE_TYPENAME(int, "int32")
MFX_READABLE_TYPENAME(unsigned int, "uint32")
MFX_READABLE_TYPENAME(long long, "int64")
MFX_READABLE_TYPENAME(unsigned long long, "uint64")
MFX_READABLE_TYPENAME(float, "fp32")
MFX_READABLE_TYPENAME(double, "fp64")

#undef MFX_READABLE_TYPENAME
