"""Writes tests/golden/known_answers.json: the known answers stated inline in the
reference's own tests (inputs and expected outputs exactly as the reference
asserts them, with the file:line of each), plus the input corpora those tests
compare GPU against CPU on.  Nothing here runs or imports the reference: the
values were read from its test sources and are restated as data.
"""
import json
import os

NONE = 0xFFFFFFFF
K = {"blank": 0, "text": 1, "heading": 2, "element": 3, "list": 4, "connect": 5, "ordered": 6, "comment": 7}

data = {
    "scan_line_offsets": [
        {"src": "crates/oneline-simd/src/lib.rs:530-534", "text": "line1\nline2\r\nline3\n", "offsets": [0, 6, 13]},
        {"src": "crates/oneline-gpu/src/tests.rs:75-83",
         "text": "|=[ Header ]\n| Body text\n||:rule(x)\n|| Content.\n", "offsets": [0, 13, 25, 36]},
    ],
    "parse_line": [
        {"src": "crates/oneline-simd/src/lib.rs:538-540", "line": "| @card(title:\"test\")\n", "depth": 1, "kind": K["element"]},
        {"src": "crates/oneline-simd/src/lib.rs:542-544", "line": "||:rule(allow:list(card))\n", "depth": 2, "kind": K["connect"]},
        {"src": "crates/oneline-simd/src/lib.rs:546-548", "line": "|||- List item\n", "depth": 3, "kind": K["list"]},
        {"src": "crates/oneline-simd/src/lib.rs:550-552", "line": "|-. Ordered\n", "depth": 1, "kind": K["ordered"]},
        {"src": "crates/oneline-simd/src/lib.rs:553", "line": "-.x\n", "kind": K["text"]},
        {"src": "crates/oneline-simd/src/lib.rs:554", "line": "|| // c\n", "kind": K["comment"]},
        {"src": "crates/oneline-simd/src/lib.rs:555", "line": "/x\n", "kind": K["text"]},
        {"src": "crates/oneline-simd/src/lib.rs:557-559", "line": "|=[ Header ]\n", "depth": 1, "kind": K["heading"]},
    ],
    "parse_document": [
        {"src": "crates/oneline-simd/src/lib.rs:564-575",
         "text": "|=[ Header ]\n| Is this syntax necessary?\n||:rule(x)\n|| Body text.\n",
         "depth_kind": [[1, K["heading"]], [1, K["text"]], [2, K["connect"]], [2, K["text"]]]},
        {"src": "crates/oneline-gpu/src/tests.rs:28-50",
         "text": "|=[ Header ]\n| Body text\n||:rule(x)\n|| Content.\n",
         "depth_kind": [[1, K["heading"]], [1, K["text"]], [2, K["connect"]], [2, K["text"]]]},
        {"src": "crates/oneline-gpu/src/tests.rs:86-105 (Element lines 0,2; depth-2 lines 1,3)",
         "text": "| @card(title:\"A\")\n|| Content\n| @card(title:\"B\")\n|| More\n|=[ Header ]\n",
         "depth_kind": [[1, K["element"]], [2, K["text"]], [1, K["element"]], [2, K["text"]], [1, K["heading"]]]},
    ],
    "tree_depths": [
        {"src": "crates/oneline-core/src/links.rs:92-111",
         "lines": [[1, False], [2, False], [3, False], [2, False], [0, True], [1, False], [0, False]],
         "parent": [NONE, 0, 1, 0, NONE, NONE, NONE], "subtree_end": [5, 3, 3, 5, 5, 6, 7]},
    ],
    "tree_text": [
        {"src": "crates/oneline-simd/src/links.rs:305-313", "text": "|=[ A ]\n|| x\n||- y\n|||- z\n|=[ B ]\n|| w\n",
         "subtree": [[0, 0, 4], [4, 4, 6]], "parent_of": [[3, 2]]},
    ],
    "find_all": [
        {"src": "crates/oneline-simd/src/search.rs:131", "text": "aaaa", "pattern": "aa", "offsets": [0, 1, 2]},
        {"src": "crates/oneline-simd/src/search.rs:132", "text": "abcabc", "pattern": "abc", "offsets": [0, 3]},
        {"src": "crates/oneline-simd/src/search.rs:133", "text": "abc", "pattern": "abcd", "offsets": []},
        {"src": "crates/oneline-simd/src/search.rs:134", "text": "", "pattern": "a", "offsets": []},
        {"src": "crates/oneline-simd/src/search.rs:135", "text": "xabc", "pattern": "abc", "offsets": [1]},
        {"src": "crates/oneline-simd/src/search.rs:136", "text": "a\nb", "pattern": "a\nb", "offsets": [0]},
        {"src": "crates/oneline-simd/src/search.rs:127", "text": "aaaa", "pattern": "aa", "count": 3},
    ],
    # Inputs on which the reference asserts GPU == CPU (expected values come from the CPU semantics).
    "parse_edge_corpus": {
        "src": "crates/oneline-gpu/src/tests.rs:114-156 and crates/oneline-simd/src/lib.rs:578-592",
        "cases": ["x", "x\n", "\n", "\r\n", "a\r\nb", "||||\n|||| \n", "```\n- x\n```\n", "| 日本語\n",
                  "-\tx\n", "- x\n", "-x\n", "-\n", "- \n", ":rule(a)\n|| :{y:2}\n", "=\n=[ t ]\n", "@\n",
                  "> a\n* b\n1. c\n", "|-\tx\n|:y\n", "// c\n", "//\n", "/x\n", "/\n", "|| // c\n", "http://x\n",
                  "/// doc\n", "| | x\n", "|  | x\n", "|\t|x\n", "| x | y\n", "| |\n", "|| |\n", " | x\n",
                  "|| \t | \n", "-. x\n", "-.\tx\n", "-.x\n", "-.\n", "-. \n", "-..\n", "- -. x\n",
                  "", "\n\n\n", "x\ny", "a\r\nb\r\n", "|\n", "   \n\t\n", "| 日本語\n|| ｗｉｄｅ\n",
                  "```\n- not a list\n```\n", "no trailing newline"],
    },
    "scan_corpus": {
        "src": "crates/oneline-gpu/src/tests.rs:54-72",
        "cases": ["hello", "hello\n", "line1\nline2\nline3\n", "line1\r\nline2\r\nline3\r\n", "\n\n\n",
                  "|=[ Header ]\n| Paragraph\n|| Content\n"],
    },
    "search_corpus": {
        "src": "crates/oneline-gpu/src/tests.rs:321-338 (alignment, overlap, tail, NUL padding)",
        "cases": [["aaaaaaaa", "aa"], ["abc", "abc"], ["xxabc", "abc"], ["ab", "abc"], ["a\nb\na\nb", "a\nb"],
                  ["xa", "a\u0000"], ["xa\u0000", "a\u0000"]] +
                 [["x" * s + "needle" + "y" * s + "needle", "needle"] for s in range(8)],
    },
    "search_rejects": {
        "src": "crates/oneline-gpu/src/tests.rs:340-357", "text": "abc",
        "bad_lengths": [0, 65], "ok_length": 64, "after_error": {"pattern": "b", "offsets": [1]},
    },
    "many_corpus": {
        "src": "crates/oneline-gpu/src/tests.rs:805-826 (overlap, nesting, duplicate ids, NUL, edges)",
        "text": "abcdefabcdefabcd\u0000\u0000\u0000\u0000abcdxabcdef|||||- - -\nabcdabcd",
        "patterns": ["abcd", "abcdef", "bcde", "abcdef", "cdefabcd", "abcdabcd", "\u0000\u0000\u0000\u0000", "||||",
                     "- - -", "\n|", "abcdx", "xxxx-not-there", "abcdabcd"],
        "also": [["abcdefgh", ["abcdefgh", "abcd", "efgh"]], ["abcdefgh", ["abcdefghi", "zzzz"]]],
    },
    "tree_depth_limit": {
        "src": "crates/oneline-gpu/src/tests.rs:570-591",
        "ok": "|" * 31 + " x\n",
        "too_deep": "| a\n" + "|" * 32 + " x\n| b\n",
        "blank_deep_ok": "| a\n" + "|" * 40 + "\n| b\n",
        "error_mentions": "32",
    },
}

here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "known_answers.json"), "w") as f:
    json.dump(data, f, indent=1, ensure_ascii=True)
print("wrote known_answers.json")
