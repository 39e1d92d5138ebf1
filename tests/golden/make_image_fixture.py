"""Writes tests/golden/G6a_XZ_128.npy: the binary microstructure image the reference's
image-to-dump notebook reads (benchmark/AM_Ti64/G6a_XZ_128.tif, 128 x 128, values 0/255), as a
uint8 array, so that tests run where the reference tree is absent. Run once in a container that
has the reference checkout and PIL."""
import os
import sys

import numpy as np

SRC = "/root/reference/benchmark/AM_Ti64/G6a_XZ_128.tif"

if __name__ == "__main__":
    from PIL import Image

    img = np.array(Image.open(sys.argv[1] if len(sys.argv) > 1 else SRC))
    assert img.dtype == np.uint8 and set(np.unique(img)) <= {0, 255}
    np.save(os.path.join(os.path.dirname(os.path.abspath(__file__)), "G6a_XZ_128.npy"), img)
    print(img.shape, int((img == 255).sum()))
