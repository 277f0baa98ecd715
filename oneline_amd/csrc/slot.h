// slot.h -- serving slots (host side, used by serve.cpp): a context plus one
// device arena that every batch of the slot is carved from, so running a
// batch allocates no device memory.  Not part of the public C ABI.
#pragma once

#include "../../include/oneline_hip.h"

#include <cstddef>
#include <cstdint>
#include <vector>

struct ol_slot;

// Results of one document of a slot batch (offsets relative to the document).
struct SlotResult {
    std::vector<uint32_t> offsets, parent, se, fc, ns, m_off, m_id;
    std::vector<uint8_t> classes;
};

// A slot with `arena_bytes` of device memory (allocated once, here).
int ol_slot_create(size_t arena_bytes, ol_slot **out);
void ol_slot_destroy(ol_slot *s);
ol_ctx *ol_slot_ctx(ol_slot *s);
size_t ol_slot_arena_bytes(const ol_slot *s);

// Arena bytes a batch of documents of these lengths takes: every per-line
// array sized for one line per text byte (a bound), and for a search the
// candidate store; a search's match list comes on top (slot_match_bytes).
size_t ol_slot_parse_bytes(const size_t *lens, size_t count);
size_t ol_slot_search_bytes(const size_t *lens, size_t count);
constexpr size_t ol_slot_match_bytes(uint64_t matches) { return (size_t)matches * 12 + 512; }

// Parse + tree of `count` documents as one batch; results[i] gets document
// i's offsets, classes and links.  One read-back (the line counts) between
// the scan and the tree.
int ol_slot_parse(ol_slot *s, const uint8_t *const *docs, const size_t *lens, size_t count, SlotResult *results);
// Dictionary search of `count` documents as one batch; results[i] gets
// document i's matches sorted by (offset, id).  One read-back (the total)
// between the count and the writes; OL_ERR_NOMEM if the match list does not
// fit what is left of the arena.
int ol_slot_search(ol_slot *s, const uint8_t *const *docs, const size_t *lens, size_t count, const ol_dict *dict,
                   SlotResult *results);
// Dictionary search of a resident document (its text already in HBM).
int ol_slot_search_doc(ol_slot *s, const ol_doc *doc, const ol_dict *dict, SlotResult *result);
// Device bytes a resident document and its tree hold.
size_t ol_doc_device_bytes(const ol_doc *d);
size_t ol_tree_device_bytes(const ol_tree *t);
// Device bytes of a compiled dictionary.
size_t ol_dict_device_bytes(const ol_dict *d);
