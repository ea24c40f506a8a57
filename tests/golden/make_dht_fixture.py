"""Extracts the Bessel-zero table that the reference ships as data
(Codes/sub_code/hankel_transform/dht.mat: first 4097 zeros of J0..J4, float64)
into dht_bessel_zeros.npz.  Run in a container that has /root/reference."""
import os

import numpy as np
import scipy.io as sio

SRC = '/root/reference/Codes/sub_code/hankel_transform/dht.mat'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'dht_bessel_zeros.npz')

if __name__ == '__main__':
    c = sio.loadmat(SRC)['c']
    np.savez_compressed(OUT, c=c)
    print('wrote', OUT, c.shape)
