// Compiled pattern index: the reference's WPSIDX01 on-disk format
// (core/src/compiled_index/) read and written on the host.
//
// Versions 1-3 are the reference's own files; they are read and rebuilt from
// the pattern sources they store, as CompiledPatternIndex::to_pattern_set
// does (core/src/compiled_index/mod.rs:287-372). Version 3 is also written,
// for sets without regexes (the reference's regex sections hold
// regex-automata DFAs, which this library does not produce).
//
// Version 5 is this library's: the same header and sections as version 3,
// then one section per regex holding its anchored DFA (four start states, an
// end-of-text column) and one holding the factor automaton, so that loading
// skips the two expensive compilation steps. Version 4, an earlier layout of
// this library's regex sections, is read and rebuilt from its stored sources.
// Every length, count and state index in a file is checked before it is used,
// and versions 3 to 5 carry a CRC-32 (IEEE) trailer.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "patterns.hpp"

namespace ws {

constexpr uint32_t kIndexVersionRef = 3;   // the reference's current version
constexpr uint32_t kIndexVersion = 5;      // this library's version
constexpr uint32_t kIndexVersionOld = 4;   // an earlier layout of this library's (rebuilt from sources)
constexpr uint32_t kIndexHashWindow = 8;   // compiler.rs:118
constexpr uint64_t kIndexMaxFileBytes = 512ull << 20;  // mod.rs:19

struct IndexHeader {
  uint32_t version = 0;
  bool ci = false;
  uint32_t pattern_count = 0, literal_count = 0, regex_count = 0, hash_window_len = 0;
  bool prebuilt = false;  // version 5: DFA and automaton tables present
};

// A parsed index: the pattern list as to_pattern_set rebuilds it and, for
// version 5, the stored tables.
struct IndexContents {
  IndexHeader hdr;
  std::vector<PatternSrc> src;
  Prebuilt pre;
};

// CRC-32 with the IEEE 802.3 polynomial (reflected 0xEDB88320, initial and
// final inversion), as builder.rs:65-78 computes it.
uint32_t crc32_ieee(const uint8_t* p, size_t n);

// Serializes a compiled set (CompiledPatternIndex::build, builder.rs:8-61).
// version is kIndexVersionRef or kIndexVersion. Throws CompileError.
std::vector<uint8_t> index_serialize(const PatternsHost& h, uint32_t version);

// Parses and checks an index (CompiledPatternIndex::parse, mod.rs:96-271)
// and rebuilds its pattern list (to_pattern_set, mod.rs:287-372). Throws
// CompileError (WS_ERR_PATTERN_COMPILATION_FAILED) on any inconsistency.
IndexContents index_parse(const uint8_t* data, size_t len);

}  // namespace ws
