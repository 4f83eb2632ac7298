"""aurora_pic: Python front end of the MI355X-native AuroraPIC particle loop.

The product is the HIP library ``libaurora_hip.so`` (C ABI declared in
``include/aurora_hip.h``) and the C++20 host classes in ``host/``; this
package only loads the library (``hip``) and builds the synthetic workloads
(``workloads``) used by the tests and the benchmark.
"""
