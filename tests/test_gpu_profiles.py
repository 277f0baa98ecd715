"""The reference's parity loops over its six generated document profiles
(Profile::ALL at 3 MiB: crates/oneline-gpu/src/tests.rs:161-168 parse,
303-329 search, 498-504 tree; crates/oneline-simd/src/search.rs:290-312
dictionary), on the HIP path against the oracle, with the profiles restated
in oneline_amd/csrc/profiles.c."""
import pytest

from oneline_amd import datagen

from test_gpu_many import assert_many_matches
from test_gpu_parse import assert_parse_matches
from test_gpu_search import assert_search_matches
from test_gpu_tree import assert_tree_matches

pytestmark = pytest.mark.gpu

SIZE = 3 << 20


@pytest.mark.parametrize("name", datagen.PROFILES)
def test_parse_matches_oracle(parser, name):
    text = datagen.profile(name, SIZE, 11)
    print(f"[gpu] parse profile {name}, {text.size} bytes", flush=True)
    assert_parse_matches(parser, text, name)


@pytest.mark.parametrize("name", datagen.PROFILES)
def test_search_matches_oracle(parser, name):
    text = datagen.profile(name, SIZE, 17)
    print(f"[gpu] search profile {name}, {text.size} bytes", flush=True)
    doc = parser.upload(text, nodes=False)
    pats = [p.encode() for p in ("lorem", "SIMD GPU", "naïve", "|", "||", "\n|",
                                 "階層 解析", "- ", "zzzz-absent")]
    # slices of the text itself, 1 byte up to the 64-byte maximum
    for i, ln in enumerate((1, 2, 3, 5, 8, 17, 33, 63, 64)):
        at = 1000 + i * 99_991
        pats.append(bytes(text[at:at + ln]))
    for p in pats:
        assert_search_matches(parser, text, p, doc)


@pytest.mark.parametrize("name", datagen.PROFILES)
def test_tree_matches_oracle(parser, name):
    text = datagen.profile(name, SIZE, 21)
    print(f"[gpu] tree profile {name}, {text.size} bytes", flush=True)
    assert_tree_matches(parser, text, name)


def test_prose_dictionary_matches_single_searches(parser):
    text = datagen.profile("prose", SIZE, 8)
    v = datagen.prose_vocabulary()
    pats = [v[i] for i in (0, 1, 5, 200, 3500, 4000)] + [b"|", b"- "]
    assert_many_matches(parser, text, pats, "prose")
    for p in pats:
        assert_search_matches(parser, text, p)
