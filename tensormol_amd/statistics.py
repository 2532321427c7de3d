"""Running statistics used by the drivers.

OnlineEstimator has the interface of TensorMol/Math/Statistics.py:8-27: it is
built from a first sample and each call with a new sample returns the running
mean and the sample variance of everything seen.  It follows Welford's
single-pass update (Welford 1962; Knuth, TAOCP vol. 2, 4.2.2): with the
count n, mean and m2 = sum of squared deviations, a new x gives
mean += (x - mean) / n and m2 += (x - mean_old)(x - mean_new).
"""


class OnlineEstimator:
    def __init__(self, x_):
        self.n = 1
        self.mean = x_ * 1.0
        self.m2 = x_ * 0.0

    def __call__(self, x_):
        self.n += 1
        before = x_ - self.mean
        self.mean = self.mean + before / self.n
        self.m2 = self.m2 + before * (x_ - self.mean)
        return self.mean, self.m2 / (self.n - 1)
