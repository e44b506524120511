/*
 * Text rules of the reference's readers (Rust std): UTF-8 validation, as
 * BufRead::lines() applies it to every line, and str::trim, which strips
 * leading and trailing code points with the Unicode White_Space property.
 * Header-only so that the library and the CLI share one definition.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace dto {

/* Decodes the code point starting at s[i] of a valid UTF-8 string; returns
 * its byte length through *len. */
static inline uint32_t utf8_at(const std::string &s, size_t i, size_t *len) {
  const unsigned char c = (unsigned char)s[i];
  if (c < 0x80) {
    *len = 1;
    return c;
  }
  if (c < 0xE0) {
    *len = 2;
    return ((c & 0x1Fu) << 6) | ((unsigned char)s[i + 1] & 0x3Fu);
  }
  if (c < 0xF0) {
    *len = 3;
    return ((c & 0x0Fu) << 12) | (((unsigned char)s[i + 1] & 0x3Fu) << 6) | ((unsigned char)s[i + 2] & 0x3Fu);
  }
  *len = 4;
  return ((c & 0x07u) << 18) | (((unsigned char)s[i + 1] & 0x3Fu) << 12) | (((unsigned char)s[i + 2] & 0x3Fu) << 6) |
         ((unsigned char)s[i + 3] & 0x3Fu);
}

/* Well-formed UTF-8 (RFC 3629: shortest forms only, no surrogates, at most
 * U+10FFFF), the check str::from_utf8 makes. */
static inline bool utf8_valid(const std::string &s) {
  const size_t n = s.size();
  size_t i = 0;
  while (i < n) {
    const unsigned char c = (unsigned char)s[i];
    if (c < 0x80) {
      ++i;
      continue;
    }
    size_t len;
    unsigned char lo = 0x80, hi = 0xBF; /* allowed range of the second byte */
    if (c >= 0xC2 && c <= 0xDF) {
      len = 2;
    } else if (c >= 0xE0 && c <= 0xEF) {
      len = 3;
      if (c == 0xE0) lo = 0xA0;
      if (c == 0xED) hi = 0x9F; /* no surrogates */
    } else if (c >= 0xF0 && c <= 0xF4) {
      len = 4;
      if (c == 0xF0) lo = 0x90;
      if (c == 0xF4) hi = 0x8F;
    } else {
      return false;
    }
    if (i + len > n) return false;
    const unsigned char c1 = (unsigned char)s[i + 1];
    if (c1 < lo || c1 > hi) return false;
    for (size_t t = 2; t < len; ++t)
      if (((unsigned char)s[i + t] & 0xC0u) != 0x80u) return false;
    i += len;
  }
  return true;
}

/* Unicode White_Space (PropList.txt): the code points str::trim removes. */
static inline bool unicode_white_space(uint32_t cp) {
  return (cp >= 0x09 && cp <= 0x0D) || cp == 0x20 || cp == 0x85 || cp == 0xA0 || cp == 0x1680 ||
         (cp >= 0x2000 && cp <= 0x200A) || cp == 0x2028 || cp == 0x2029 || cp == 0x202F || cp == 0x205F ||
         cp == 0x3000;
}

/* str::trim of a valid UTF-8 string. */
static inline std::string rust_trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b) {
    size_t len;
    const uint32_t cp = utf8_at(s, a, &len);
    if (!unicode_white_space(cp)) break;
    a += len;
  }
  while (b > a) {
    size_t st = b - 1; /* start of the last code point */
    while (st > a && ((unsigned char)s[st] & 0xC0u) == 0x80u) --st;
    size_t len;
    const uint32_t cp = utf8_at(s, st, &len);
    if (!unicode_white_space(cp)) break;
    b = st;
  }
  return s.substr(a, b - a);
}

}  // namespace dto
