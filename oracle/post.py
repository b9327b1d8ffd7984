"""Pure-Python restatement of the reference's post-processing and output
formats, for small cases.

TEST INFRASTRUCTURE ONLY (the checker of the library's host functions).
Follows GPU-AS/Image.cpp:1579-1778 (check_dead_seeds, label_replace.gs,
sequence_label) and :1780-1806 (print_sv_txt).
"""


def sequence_label(labels, seq, dead_seed):
    """GPU-AS/Image.cpp:1727-1778 on a flat list (modified in place)."""
    if dead_seed:
        last_label = labels[0]                 # never updated afterwards (:1735)
        last_map = seq.setdefault(last_label, 0)
        labels[0] = last_map
        for i in range(1, len(labels)):
            if labels[i] != last_label:
                if labels[i] != -1:
                    last_map = seq.setdefault(labels[i], 0)
            labels[i] = last_map
        return
    last_label = 0
    if labels[0] >= 0:
        last_label = labels[0]
    for i in range(1, len(labels)):
        if labels[i] != last_label:
            if labels[i] == -1:
                labels[i] = last_label
            last_label = labels[i]


def check_dead_seeds(counts, A, labels):
    """counts: per-slot voxel counts (A+1, slot A = label -1); labels: flat
    list in GL raster order.  Returns (labels, dead, final_label_num)."""
    labels = list(labels)
    invalid = counts[A] > 0                                 # :1590
    dead_list, seq, valid = [], {}, 0
    for i in range(A):
        if counts[i] < 1:
            dead_list.append(i)
        else:
            seq[i] = valid
            valid += 1
    if len(dead_list) == 0:
        if invalid:
            sequence_label(labels, seq, False)
    elif len(dead_list) <= 3:
        last = A
        for d in dead_list:                                 # label_replace.gs:26-29
            last -= 1
            labels = [d if l == last else l for l in labels]
        if invalid:
            sequence_label(labels, seq, False)
    else:
        sequence_label(labels, seq, True)
    return labels, len(dead_list), valid


def label_txt(labels3d):
    """GPU-AS/Image.cpp:1780-1806 text for a [D][H][W] nested list."""
    out = []
    for frame in labels3d:
        for row in reversed(frame):
            out.append("".join(f"{v:<4d} " for v in row) + "\n")
        out.append("\n\n")
    return "".join(out)
