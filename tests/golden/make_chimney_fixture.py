"""Copies the reference's chimney frame sequence (100 observed 8-bit frames,
the ground truth and the reference's own result) into tests/golden/chimney/.

The files are data (the reference's only fixture), copied byte for byte; run
in the build container, where the reference is mounted:

    python tests/golden/make_chimney_fixture.py /root/reference/chimney
"""
import os
import shutil
import sys


def main(src):
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'chimney')
    os.makedirs(dst, exist_ok=True)
    names = ['groundtruth.png', 'result.png'] + ['y_%04d.png' % i for i in range(1, 101)]
    for n in names:
        shutil.copyfile(os.path.join(src, n), os.path.join(dst, n))
    print(f'{len(names)} files -> {dst}')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/chimney')
