// host_common.hpp -- host-side helpers shared by the C-ABI library's C++ code.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ngia2.h"

namespace ngia {

// Per-thread error text for calls made without a context.
void set_thread_error(const std::string& msg);
const char* thread_error();

// Residue code table: the reduced 15-letter alphabet of makedb.hpp:161-171.
struct CodeTable {
  uint8_t t[256];
  CodeTable();
};
extern const CodeTable kCodes;

// One FASTA record as makeIndex sees it (makedb.hpp:94-127).
struct SeqIndex {
  uint64_t fasta_off;   // byte offset of the name line
  uint64_t pack_off;    // absolute byte offset of the packed record in the DB
  uint64_t text_off;    // absolute byte offset of the text record in the DB
  uint32_t name_len;    // name line length (with '>')
  uint32_t read_len;    // residues
  uint32_t seq_len;     // bytes of the record in the FASTA file
};

// FASTA checks of MakeDB::check (makedb.hpp:59-91): 0 ok, else message.
int fasta_check(const char* path, std::string* why);
// makeIndex: scan, drop reads >= 0xFFFF, stable length-descending sort,
// offsets for a 64-band DB.  total = records scanned.
int fasta_index(const char* path, std::vector<SeqIndex>* idx, uint64_t* total);
// Layout constants of the packed DB for n sequences.
inline uint64_t db_sign_base(uint64_t n) { return 4 + 24 * n; }
inline uint64_t db_pack_base(uint64_t n) { return 4 + 24 * n + 4ull * NGIA_BANDS * n; }
inline uint32_t rec_words(uint32_t L) { return 1 + (L + 31) / 32 * 4; }
// Assign pack/text offsets to a sorted index (makedb.hpp:137-148); returns file size.
uint64_t assign_offsets(std::vector<SeqIndex>* idx);
// Pack one residue string into a record (makedb.hpp:222-231).
void pack_record(const uint8_t* res, uint32_t L, uint32_t* rec);
// Records of residue strings first..first+count-1 (string q at res + res_off[q],
// res_len[q] residues), record j at dst + word_pos[j] - word_base; OpenMP.
void pack_records(const uint8_t* res, const uint64_t* res_off, const uint32_t* res_len,
                  uint32_t first, uint32_t count, const uint64_t* word_pos, uint64_t word_base,
                  uint32_t* dst);

}  // namespace ngia
