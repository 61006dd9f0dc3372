"""The reference's twelve behavioural tests (tests/test_cuckoo_filter.cu:32-344),
written once against a minimal filter interface so the same checks run on the
CPU oracle and on the HIP library.

Fixture (test_cuckoo_filter.cu:10-30): CuckooConfig<uint32_t, 16, 500, 256,
16> -- uint32 keys, 16-bit tags, DFS eviction, uint64 words; random keys are
uniform in [1, 2^32-1] from a generator seeded with 42.

An adapter provides: new(capacity) -> handle, insert(h, keys) -> occupied,
contains(h, keys) -> bool array, delete(h, keys) -> (occupied, ok array),
load_factor(h), capacity(h), clear(h).
"""
import numpy as np

TARGET_LOAD_FACTOR = 0.95


def random_keys(n, seed=42):
    rng = np.random.Generator(np.random.MT19937(seed))
    return rng.integers(1, 2**32 - 1, size=n, dtype=np.uint64, endpoint=True).astype(np.uint32)


def basic_insert_and_query(A):  # :32-51
    f = A.new(10000)
    keys = np.array([1, 2, 3, 4, 5, 100, 200, 300], np.uint32)
    assert A.insert(f, keys) == keys.size
    assert A.contains(f, keys).all()


def empty_filter(A):  # :53-69
    f = A.new(10000)
    assert A.contains(f, np.array([1, 2, 3, 4, 5], np.uint32)).sum() == 0


def clear_operation(A):  # :71-99
    f = A.new(10000)
    keys = np.array([1, 2, 3, 4, 5], np.uint32)
    A.insert(f, keys)
    assert A.contains(f, keys).all()
    A.clear(f)
    assert A.contains(f, keys).sum() == 0
    assert A.load_factor(f) == 0.0


def load_factor(A):  # :101-124
    f = A.new(10000)
    assert A.load_factor(f) == 0.0
    keys = random_keys(5000)
    inserted = A.insert(f, keys)
    lf = A.load_factor(f)
    assert 0.0 < lf <= 1.0
    assert np.float32(lf) == np.float32(np.float32(inserted) / np.float32(A.capacity(f)))
    assert inserted > 0


def near_capacity_insertion(A):  # :126-145
    n = int((1 << 20) * TARGET_LOAD_FACTOR)
    f = A.new(n)
    keys = random_keys(n)
    inserted = A.insert(f, keys)
    assert inserted > n * 0.99
    found = int(A.contains(f, keys).sum())
    assert found > inserted * 0.99


def duplicate_insertions(A):  # :147-170
    f = A.new(10000)
    keys = np.array([1, 1, 2, 2, 3, 3, 4, 4, 5, 5], np.uint32)
    assert A.insert(f, keys) == keys.size
    assert A.contains(f, np.array([1, 2, 3, 4, 5], np.uint32)).all()


def basic_deletion(A):  # :172-208
    f = A.new(10000)
    keys = np.array([1, 2, 3, 4, 5, 100, 200, 300], np.uint32)
    assert A.insert(f, keys) == keys.size
    assert A.contains(f, keys).all()
    remaining, ok = A.delete(f, keys)
    assert ok.all()
    assert remaining == 0
    assert A.contains(f, keys).sum() == 0


def delete_non_existent_keys(A):  # :210-225
    f = A.new(10000)
    _, ok = A.delete(f, np.array([1, 2, 3, 4, 5], np.uint32))
    assert ok.sum() == 0


def partial_deletion(A):  # :227-273
    f = A.new(10000)
    insert_keys = np.arange(1, 11, dtype=np.uint32)
    assert A.insert(f, insert_keys) == 10
    delete_keys = np.array([2, 4, 6, 8, 10], np.uint32)
    remaining, ok = A.delete(f, delete_keys)
    assert ok.all()
    assert remaining == 5
    assert A.contains(f, delete_keys).sum() == 0
    assert A.contains(f, np.array([1, 3, 5, 7, 9], np.uint32)).all()


def delete_and_reinsert(A):  # :275-302
    f = A.new(10000)
    keys = np.array([1, 2, 3, 4, 5], np.uint32)
    A.insert(f, keys)
    A.delete(f, keys)
    assert A.contains(f, keys).sum() == 0
    assert A.insert(f, keys) == keys.size
    assert A.contains(f, keys).all()


def load_factor_after_deletion(A):  # :304-322
    f = A.new(10000)
    keys = random_keys(5000)
    A.insert(f, keys)
    before = A.load_factor(f)
    assert before > 0.0
    remaining, _ = A.delete(f, keys[:2500])
    after = A.load_factor(f)
    assert after < before
    assert np.float32(after) == np.float32(np.float32(remaining) / np.float32(A.capacity(f)))


def delete_duplicates(A):  # :324-344
    f = A.new(10000)
    A.insert(f, np.array([1, 1, 1, 2, 2, 3], np.uint32))
    _, ok = A.delete(f, np.array([1, 1, 1], np.uint32))
    assert ok.sum() >= 1


ALL = [basic_insert_and_query, empty_filter, clear_operation, load_factor, near_capacity_insertion,
       duplicate_insertions, basic_deletion, delete_non_existent_keys, partial_deletion, delete_and_reinsert,
       load_factor_after_deletion, delete_duplicates]
