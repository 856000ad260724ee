"""Reading the FastLoad matrix file in tests: the section table and the section hash, restated
from the file layout (192-byte header with the section count at byte 120; 32-byte section entries
tag, element size, count, offset, hash)."""
import struct

M64 = (1 << 64) - 1


def section_hash(b: bytes) -> int:
    """64-bit multiply-rotate mix over 8-byte little-endian words, the tail zero-padded."""
    h = (0x9e3779b97f4a7c15 ^ len(b)) & M64
    n8 = len(b) // 8 * 8
    for i in range(0, n8, 8):
        w = int.from_bytes(b[i:i + 8], "little")
        h ^= (w * 0xff51afd7ed558ccd) & M64
        h = ((((h << 31) | (h >> 33)) & M64) * 0xc4ceb9fe1a85ec53) & M64
    w = int.from_bytes(b[n8:], "little")
    h ^= (w * 0xff51afd7ed558ccd) & M64
    h ^= h >> 33
    h = (h * 0xc4ceb9fe1a85ec53) & M64
    return h ^ (h >> 29)


def sections(raw):
    """{tag: (entry offset, element size, count, data offset, hash)}"""
    nsec = struct.unpack_from("<I", raw, 120)[0]
    out = {}
    for i in range(nsec):
        ent = 192 + 32 * i
        tag, elem, count, off, h = struct.unpack_from("<IIQQQ", raw, ent)
        out[tag] = (ent, elem, count, off, h)
    return out


def tamper(raw, tag, elem_index, value, fmt):
    """Writes `value` into element `elem_index` of section `tag` and re-hashes the section, so only
    the library's own index checks can refuse the file."""
    raw = bytearray(raw)
    ent, elem, count, off, _ = sections(raw)[tag]
    struct.pack_into("<" + fmt, raw, off + elem * elem_index, value)
    struct.pack_into("<Q", raw, ent + 24, section_hash(bytes(raw[off:off + elem * count])))
    return raw
