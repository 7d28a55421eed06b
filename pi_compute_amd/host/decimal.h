// decimal.h - parallel divide-and-conquer conversion of a non-negative mpz to
// decimal (the role of the reference's BaseConverter::parallel_to_decimal,
// src/io/base_converter.h:37-41; same output: the digits of n, no sign).
//
// Powers 10^(2^i) are built once by squaring; a number known to hold at most
// `width` digits is split at the largest power below width, the halves are
// converted independently (threads at the top levels) and written into their
// slots of one preallocated string, left-padded with '0'.
#pragma once

#include <gmp.h>

#include <cstddef>
#include <string>

namespace pi {

std::string to_decimal(mpz_srcptr n, unsigned threads = 0);

}  // namespace pi
